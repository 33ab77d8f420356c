// ietl/tmatrix.h -- the Lanczos T matrix and its analysis.
//
// Same interface and semantics as ietl::Tmatrix of the reference
// (src/ietl/tmatrix.h:24-228): push_back of (alpha, beta); eigenvalues(),
// errors(), multiplicities() returning distinct eigenvalues of T (values
// closer than thold merged), with Cullum-Willoughby ghosts removed (an
// eigenvalue of multiplicity one that is also an eigenvalue, within multol,
// of T without its first row and column); error bound |beta_n z(n-1,i)| for
// multiplicity one, 0 otherwise.  Tolerances:
//   multol = max(alpha_max, beta_max) * 2 eps * (10000 + n)
//   thold  = max(eps^(2/3) * max(lambda_0, lambda_{n-1}), 5 multol)
// (the inner max is the signed one, as in the reference).
// The eigensolver is ietl/tridiag.h (no LAPACK).
#ifndef ALPS_ED_IETL_TMATRIX_H
#define ALPS_ED_IETL_TMATRIX_H

#include <algorithm>
#include <cmath>
#include <limits>
#include <stdexcept>
#include <utility>
#include <vector>

#include <ietl/traits.h>
#include <ietl/tridiag.h>

namespace ietl {

// Analysis of a tridiagonal given by alpha[0..n) and beta[0..n) (beta[n-1] is
// the coupling to the next Lanczos vector), independent of any vector space.
struct tmatrix_analysis {
  std::vector<double> eigval_distinct, eigval_noghost, err, err_noghost;
  std::vector<double> eigval_all;  // every eigenvalue of T, ascending
  std::vector<int> mult, mult_noghost;
  // position of each distinct value's first member among the sorted
  // eigenvalues of T (selects its eigenvector, e.g. for Ritz vectors)
  std::vector<int> index_distinct, index_noghost;
  double thold = 0.0, multol = 0.0;

  void run(const double* alpha, const double* beta, int n, double alpha_max, double beta_max,
           bool do_err) {
    eigval_distinct.clear();
    eigval_noghost.clear();
    err.clear();
    err_noghost.clear();
    mult.clear();
    mult_noghost.clear();
    index_distinct.clear();
    index_noghost.clear();
    if (n <= 0) return;
    std::vector<double> eval(alpha, alpha + n);
    std::vector<double> zlast(do_err ? n : 0);
    if (tridiag::eig(n, eval.data(), beta, do_err ? zlast.data() : nullptr) > 0)
      throw std::runtime_error("LAPACK error, stev function failed.");
    eigval_all = eval;
    const double eps = std::numeric_limits<double>::epsilon();
    const double error_tol = std::pow(eps, 2. / 3.);
    multol = std::max(alpha_max, beta_max) * 2 * eps * (10000 + n);
    thold = std::max(eval[0], eval[n - 1]);
    thold = std::max(error_tol * thold, 5 * multol);
    const double bn = beta[n - 1];
    // merge eigenvalues closer than thold to the first of their group
    double group_first = eval[0];
    eigval_distinct.push_back(eval[0]);
    index_distinct.push_back(0);
    int count = 1;
    auto close_group = [&](int last_index) {
      mult.push_back(count);
      if (do_err) err.push_back(count > 1 ? 0.0 : std::fabs(bn * zlast[last_index]));
    };
    for (int i = 1; i < n; ++i) {
      if (eval[i] - group_first > thold) {
        close_group(i - 1);
        eigval_distinct.push_back(eval[i]);
        index_distinct.push_back(i);
        group_first = eval[i];
        count = 1;
      } else {
        ++count;
      }
    }
    close_group(n - 1);
    // ghosts: compare with the eigenvalues of T minus its first row/column;
    // the scan position advances monotonically across the distinct values
    std::vector<double> reduced(alpha + 1, alpha + n);
    if (n > 1 && tridiag::eig(n - 1, reduced.data(), beta + 1, nullptr) > 0)
      throw std::runtime_error("LAPACK error, stev function failed.");
    int scan = 0;
    for (std::size_t i = 0; i < eigval_distinct.size(); ++i) {
      if (mult[i] != 1) continue;
      const double lam = eigval_distinct[i];
      while (scan < n - 2) {
        const double mu = reduced[scan];
        if (mu - lam >= multol) break;
        ++scan;
        if (std::fabs(lam - mu) < multol) {
          mult[i] = 0;
          if (do_err) err[i] = 0.0;
          break;
        }
      }
    }
    for (std::size_t i = 0; i < eigval_distinct.size(); ++i) {
      if (mult[i] == 0) continue;
      eigval_noghost.push_back(eigval_distinct[i]);
      mult_noghost.push_back(mult[i]);
      index_noghost.push_back(index_distinct[i]);
      if (do_err) err_noghost.push_back(err[i]);
    }
  }
};

template <class VS>
class Tmatrix {
 public:
  typedef typename vectorspace_traits<VS>::scalar_type scalar_type;
  typedef typename vectorspace_traits<VS>::vector_type vector_type;
  typedef typename vectorspace_traits<VS>::magnitude_type magnitude_type;
  typedef typename vectorspace_traits<VS>::size_type size_type;

  Tmatrix() : error_tol(std::pow(std::numeric_limits<magnitude_type>::epsilon(), 2. / 3.)) {}

  void push_back(magnitude_type a, magnitude_type b) {
    computed_ = computed_err_ = false;
    alpha.push_back(a);
    beta.push_back(b);
    if (alpha.size() == 1) {
      alpha_max_ = a;
      beta_min_ = beta_max_ = b;
    } else {
      alpha_max_ = std::max(alpha_max_, a);
      beta_max_ = std::max(beta_max_, b);
      beta_min_ = std::min(beta_min_, b);
    }
  }
  void push_back(std::pair<magnitude_type, magnitude_type> ab) { push_back(ab.first, ab.second); }

  const std::vector<magnitude_type>& eigenvalues(bool discard_ghosts = true) const {
    if (!computed_) compute(false);
    return discard_ghosts ? an_.eigval_noghost : an_.eigval_distinct;
  }
  const std::vector<magnitude_type>& errors(bool discard_ghosts = true) const {
    if (!computed_ || !computed_err_) compute(true);
    return discard_ghosts ? an_.err_noghost : an_.err;
  }
  const std::vector<int>& multiplicities(bool discard_ghosts = true) const {
    if (!computed_) compute(false);
    return discard_ghosts ? an_.mult_noghost : an_.mult;
  }
  const std::vector<magnitude_type>& diagonal() const { return alpha; }
  const std::vector<magnitude_type>& subdiagonal() const { return beta; }

  template <class Archive>
  void save(Archive& ar) const {
    ar << alpha << beta << error_tol << alpha_max_ << beta_max_ << beta_min_;
  }
  template <class Archive>
  void load(Archive& ar) {
    ar >> alpha >> beta >> error_tol >> alpha_max_ >> beta_max_ >> beta_min_;
    computed_ = computed_err_ = false;
  }

 protected:
  std::vector<magnitude_type> alpha;
  std::vector<magnitude_type> beta;
  magnitude_type error_tol;
  magnitude_type thold_value() const { return an_.thold; }

 private:
  void compute(bool do_err) const {
    an_.run(alpha.data(), beta.data(), (int)alpha.size(), alpha_max_, beta_max_, do_err);
    computed_ = true;
    computed_err_ = do_err;
  }
  mutable tmatrix_analysis an_;
  mutable bool computed_ = false;
  mutable bool computed_err_ = false;
  magnitude_type alpha_max_ = 0, beta_max_ = 0, beta_min_ = 0;
};

}  // namespace ietl

#endif
