// ietl/iteration.h -- Lanczos iteration controls with the reference's names
// and stopping rules (src/ietl/iteration.h:67-216):
//   basic_lanczos_iteration::finished  converged -> true; else i < max_iter
//                                      -> false; else error 1, true
//   lanczos_iteration_nlowest          after >1 iterations, the lowest n
//                                      distinct non-ghost eigenvalues all
//                                      have err <= max(atol, rtol*|lambda|);
//                                      rtol = atol = 100 eps by default
//   lanczos_iteration_nhighest         the same for the highest n
//   lanczos_nlowest_better             nlowest tested every check_each steps
//   fixed_lanczos_iteration            never converged: exactly max_iter steps
#ifndef ALPS_ED_IETL_ITERATION_H
#define ALPS_ED_IETL_ITERATION_H

#include <algorithm>
#include <cmath>
#include <limits>
#include <string>
#include <vector>

namespace ietl {

template <class T>
class basic_iteration {
 public:
  basic_iteration(unsigned int max_iter, T reltol = 0., T abstol = 0.)
      : error(0), i(0), max_iter_(max_iter), rtol_(reltol), atol_(abstol) {}
  bool finished(T r, T lambda) {
    if (converged(r, lambda)) return true;
    if (i < max_iter_) return false;
    fail(1, "maximum number of iterations exceeded");
    return true;
  }
  bool converged(T r, T lambda) { return r <= rtol_ * std::fabs(lambda) || r < atol_; }
  void operator++() { ++i; }
  bool first() { return i == 0; }
  int error_code() { return error; }
  unsigned int iterations() { return i; }
  T relative_tolerance() { return rtol_; }
  T absolute_tolerance() { return atol_; }
  unsigned int max_iterations() { return max_iter_; }
  void fail(int err_code) { error = err_code; }
  void fail(int err_code, const std::string& msg) {
    error = err_code;
    err_msg = msg;
  }

 protected:
  int error;
  unsigned int i;
  unsigned int max_iter_;
  T rtol_, atol_;
  std::string err_msg;
};

template <class T, class Derived>
class basic_lanczos_iteration {
 public:
  basic_lanczos_iteration(unsigned int max_iter, T r = 0., T a = 0.)
      : error(0), i(0), max_iter_(max_iter), rtol_(r), atol_(a) {}

  template <class Tmatrix>
  bool finished(const Tmatrix& tmatrix) {
    if (static_cast<const Derived&>(*this).converged(tmatrix)) return true;
    if (i < max_iter_) return false;
    fail(1, "maximum number of iterations exceeded");
    return true;
  }
  bool converged() const { return false; }
  void operator++() { ++i; }
  bool first() const { return i == 0; }
  int error_code() const { return error; }
  unsigned int iterations() const { return i; }
  unsigned int max_iterations() const { return max_iter_; }
  T relative_tolerance() const { return rtol_; }
  T absolute_tolerance() const { return atol_; }
  void fail(int err_code) { error = err_code; }
  void fail(int err_code, const std::string& msg) {
    error = err_code;
    err_msg = msg;
  }
  const std::string& error_message() const { return err_msg; }

 protected:
  int error;
  unsigned int i;
  unsigned int max_iter_;
  T rtol_, atol_;
  std::string err_msg;
};

namespace detail {
// True when the first n (lowest) or last n (highest) distinct non-ghost
// eigenvalues satisfy err <= max(atol, rtol*|lambda|).
template <class T, class Tmatrix>
bool n_converged(const Tmatrix& tmatrix, unsigned int n, T rtol, T atol, bool lowest) {
  const std::vector<T>& errs = tmatrix.errors();
  const std::vector<T>& vals = tmatrix.eigenvalues();
  if (lowest ? vals.size() < n : errs.size() < n) return false;
  for (unsigned int k = 0; k < n; ++k) {
    const std::size_t e = lowest ? k : errs.size() - k - 1;
    const std::size_t v = lowest ? k : vals.size() - k - 1;
    if (errs[e] > std::max(atol, rtol * std::fabs(vals[v]))) return false;
  }
  return true;
}
}  // namespace detail

template <class T>
class lanczos_iteration_nlowest : public basic_lanczos_iteration<T, lanczos_iteration_nlowest<T> > {
  typedef basic_lanczos_iteration<T, lanczos_iteration_nlowest<T> > super_type;

 public:
  lanczos_iteration_nlowest(unsigned int max_iter, unsigned int n = 1,
                            T r = 100. * std::numeric_limits<T>::epsilon(),
                            T a = 100. * std::numeric_limits<T>::epsilon())
      : super_type(max_iter, r, a), n_(n) {}
  template <class Tmatrix>
  bool converged(const Tmatrix& tmatrix) const {
    if (super_type::iterations() <= 1) return false;
    return detail::n_converged<T>(tmatrix, n_, super_type::relative_tolerance(),
                                  super_type::absolute_tolerance(), true);
  }
  unsigned int n_eigenvalues() const { return n_; }

 private:
  unsigned int n_;
};

template <class T>
class lanczos_nlowest_better : public basic_lanczos_iteration<T, lanczos_nlowest_better<T> > {
  typedef basic_lanczos_iteration<T, lanczos_nlowest_better<T> > super_type;

 public:
  lanczos_nlowest_better(unsigned int max_iter, unsigned int n,
                         T r = 100. * std::numeric_limits<T>::epsilon(),
                         T a = 100. * std::numeric_limits<T>::epsilon(),
                         unsigned int check_each = 50)
      : super_type(max_iter, r, a), n_(n), ce_(check_each) {}
  template <class Tmatrix>
  bool converged(const Tmatrix& tmatrix) const {
    if (super_type::iterations() <= 1 || super_type::iterations() % ce_ != 0) return false;
    return detail::n_converged<T>(tmatrix, n_, super_type::relative_tolerance(),
                                  super_type::absolute_tolerance(), true);
  }

 private:
  unsigned int n_, ce_;
};

template <class T>
class lanczos_iteration_nhighest : public basic_lanczos_iteration<T, lanczos_iteration_nhighest<T> > {
  typedef basic_lanczos_iteration<T, lanczos_iteration_nhighest<T> > super_type;

 public:
  lanczos_iteration_nhighest(unsigned int max_iter, unsigned int n = 1,
                             T r = 100. * std::numeric_limits<T>::epsilon(),
                             T a = 100. * std::numeric_limits<T>::epsilon())
      : super_type(max_iter, r, a), n_(n) {}
  template <class Tmatrix>
  bool converged(const Tmatrix& tmatrix) const {
    if (super_type::iterations() <= 1) return false;
    return detail::n_converged<T>(tmatrix, n_, super_type::relative_tolerance(),
                                  super_type::absolute_tolerance(), false);
  }

 private:
  unsigned int n_;
};

template <class T>
class fixed_lanczos_iteration : public basic_lanczos_iteration<T, fixed_lanczos_iteration<T> > {
 public:
  explicit fixed_lanczos_iteration(unsigned int max_iter)
      : basic_lanczos_iteration<T, fixed_lanczos_iteration<T> >(max_iter, 0., 0.) {}
  template <class Tmatrix>
  bool converged(const Tmatrix&) const { return false; }
};

}  // namespace ietl

#endif
