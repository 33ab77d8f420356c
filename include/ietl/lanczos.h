// ietl/lanczos.h -- ietl::lanczos<MATRIX, VS> with the reference's interface
// and recurrence (src/ietl/lanczos.h:64-129,341-406):
//   make_first_step   v0 = gen; project; v0 /= |v0|; w = H v0;
//                     a = Re<v0,w>; w -= a v0; b = |w|; v1 = w/b
//   make_step(j)      w = H v_j; a = Re<v_j,w>; w -= a v_j;
//                     w -= beta[j-1] v_{j-1}; b = |w|; w /= b; rotate
//   generate_tmatrix  first step when T is empty, then make_step while
//                     !iter.finished(*this)
//   eigenvectors      Ritz vectors by the Cullum-Willoughby procedure of
//                     lanczos.h:134-336 (m1/m2 search, T-eigenvector at a
//                     length ma grown until its error bound is below
//                     error_tol, then a replay of the recurrence from the
//                     same generator accumulating the Ritz vectors)
// No reorthogonalization, exactly as the reference.  The only addition is a
// customization point: the step body is `lanczos_step(matrix, v_prev, v_cur,
// w, beta_prev)`, found by argument-dependent lookup, so that a device
// vector space can run the step as fused kernels (alps_ed/ietl_gpu.hpp).
// The generic version below performs the reference's operation sequence.
#ifndef ALPS_ED_IETL_LANCZOS_H
#define ALPS_ED_IETL_LANCZOS_H

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <utility>
#include <vector>

#include <ietl/iteration.h>
#include <ietl/tmatrix.h>
#include <ietl/traits.h>
#include <ietl/vectorspace.h>

namespace ietl {

template <class MATRIX, class V>
std::pair<double, double> lanczos_step(const MATRIX& matrix, V& v_prev, V& v_cur, V& w,
                                       double beta_prev) {
  mult(matrix, v_cur, w);
  const double a = real(dot(v_cur, w));
  w -= a * v_cur;
  w -= beta_prev * v_prev;
  const double b = two_norm(w);
  w /= b;
  return std::make_pair(a, b);
}

// What eigenvectors() found for each requested eigenvalue (lanczos.h:30-62):
// the lengths m1 (first T that has the value) and m2 (first T that has it
// twice), the length ma whose T-eigenvector was used, the T eigenvalue, the
// residual |H v - lambda v| and a status.
template <class magnitude_type = double>
class Info {
 public:
  enum errorinfo { ok = 0, no_eigenvalue, not_calculated };
  Info() {}
  Info(std::vector<int> M1, std::vector<int> M2, std::vector<int> Ma, std::vector<magnitude_type> Eigenvalue,
       std::vector<magnitude_type> Residuum, std::vector<errorinfo> Status)
      : m1_(M1), m2_(M2), ma_(Ma), eigenvalue_(Eigenvalue), residuum_(Residuum), status_(Status) {}
  int m1(int i) const { return m1_[i]; }
  int m2(int i) const { return m2_[i]; }
  int ma(int i) const { return ma_[i]; }
  int size() { return (int)m1_.size(); }
  magnitude_type eigenvalue(int i) const { return eigenvalue_[i]; }
  magnitude_type residual(int i) const { return residuum_[i]; }
  errorinfo error_info(int i) const { return status_[i]; }

 private:
  std::vector<int> m1_, m2_, ma_;
  std::vector<magnitude_type> eigenvalue_, residuum_;
  std::vector<errorinfo> status_;
};

template <class MATRIX, class VS>
class lanczos : public Tmatrix<VS> {
  typedef Tmatrix<VS> super_type;

 public:
  typedef typename vectorspace_traits<VS>::vector_type vector_type;
  typedef typename vectorspace_traits<VS>::scalar_type scalar_type;
  typedef typename vectorspace_traits<VS>::magnitude_type magnitude_type;

  lanczos(const MATRIX& matrix, const VS& vec)
      : matrix_(matrix), vecspace_(vec), startvector(new_vector(vec)), vec2(new_vector(vec)), n(0) {}

  template <class IT, class GEN>
  void calculate_eigenvalues(IT& iter, GEN gen) {
    generate_tmatrix(iter, gen);
  }

  template <class IT>
  void more_eigenvalues(IT& iter) {
    generate_tmatrix(iter);
  }

  // Ritz vectors of the eigenvalues [first, last) (values of eigenvalues()),
  // written to `out`; `gen` must be the generator the T matrix was built
  // with (the recurrence is replayed from it and must reproduce alpha and
  // beta exactly).
  template <class IN, class OUT, class GEN>
  void eigenvectors(IN first, IN last, OUT out, Info<magnitude_type>& inf, GEN gen, int maxiter = 0,
                    int maxcount = 50) {
    typedef typename Info<magnitude_type>::errorinfo errorinfo;
    vector_type vec3 = new_vector(vecspace_);
    super_type::eigenvalues();  // the merge threshold of the current T
    const magnitude_type thold = super_type::thold_value();
    const magnitude_type error_tol = super_type::error_tol;
    Tvectors.clear();
    std::vector<vector_type> ritz;
    const int n1 = (int)super_type::alpha.size();
    int maMax = 0;
    std::vector<magnitude_type> lambdas, residuum;
    std::vector<errorinfo> status;
    Ma.clear();
    find_m1m2(first, last);
    size_t idx = 0;
    for (IN it = first; it != last; ++it, ++idx) {
      const magnitude_type want = *it;
      magnitude_type lambda = 0;
      errorinfo st = Info<magnitude_type>::ok;
      int count = 0, ma = 0, deltam = 1;
      const int m1 = M1[idx], m2 = M2[idx];
      if (m1 != 0 && m2 != 0) {
        ma = (3 * m1 + m2) / 4 + 1;
        deltam = ((3 * m1 + 5 * m2) / 8 + 1 - ma) / 10 + 1;
      } else if (m1 != 0) {
        ma = (5 * m1) / 4 + 1;
        const magnitude_type mamax = std::min((11 * n1) / 8 + 12, (13 * m1) / 8 + 1);
        deltam = int((mamax - ma) / 10) + 1;
        if (maxiter > 0) maxcount = maxiter / deltam;
      } else {
        st = Info<magnitude_type>::no_eigenvalue;
      }
      ritz.push_back(new_vector(vecspace_));
      Tvectors.push_back(std::vector<magnitude_type>());
      if (ma == 0) ritz.back() *= 0.;
      if (ma != 0) {
        magnitude_type error = 0;
        int nth = -1;
        std::vector<magnitude_type> eval, z;
        do {
          if (ma > (int)super_type::alpha.size()) {
            fixed_lanczos_iteration<magnitude_type> grow(ma);
            generate_tmatrix(grow, gen);
          }
          ++count;
          eval.assign(super_type::alpha.begin(), super_type::alpha.begin() + ma);
          z.assign((size_t)ma * ma, 0.0);
          if (tridiag::eig_full(ma, eval.data(), super_type::beta.data(), z.data()) > 0)
            throw std::runtime_error("LAPACK error, stev function failed.");
          for (nth = ma - 1; nth >= 0; --nth)
            if (std::fabs(eval[nth] - want) <= thold) break;
          if (nth == -1) {
            error = 0;
            ma = 0;
            ritz.back() *= 0.;
            st = Info<magnitude_type>::no_eigenvalue;
          } else {
            error = std::fabs(super_type::beta[ma - 1] * z[(size_t)(ma - 1) + (size_t)nth * ma]);
            if (error > error_tol) ma += deltam;
          }
        } while (error > error_tol && count < maxcount);
        if (error > error_tol) {
          ritz.back() *= 0.;
          st = Info<magnitude_type>::not_calculated;
        } else if (ma != 0) {
          for (int i = 0; i < ma; ++i) Tvectors.back().push_back(z[(size_t)i + (size_t)nth * ma]);
          maMax = std::max(maMax, ma);
          lambda = eval[nth];
        }
      }
      lambdas.push_back(lambda);
      Ma.push_back(ma);
      status.push_back(st);
    }
    // replay the recurrence and accumulate v = sum_j z_j v_j
    std::pair<magnitude_type, magnitude_type> ab = make_first_step(gen);
    if (ab.first != super_type::alpha[0] || ab.second != super_type::beta[0])
      throw std::runtime_error("T-matrix problem at first step");
    for (size_t i = 0; i < ritz.size(); ++i)
      if (!Tvectors[i].empty()) {
        ritz[i] *= 0.;  // (handle vectors: assignment would share storage)
        ritz[i] += Tvectors[i][0] * startvector;
        ritz[i] += Tvectors[i][1] * vec2;
      }
    n = 2;
    for (int j = 2; j < maMax; ++j) {
      ab = make_step(j - 1, vec3);
      if (ab.first != super_type::alpha[j - 1] || ab.second != super_type::beta[j - 1])
        throw std::runtime_error("T-matrix problem");
      ++n;
      for (size_t i = 0; i < ritz.size(); ++i)
        if ((int)Tvectors[i].size() > j && std::fabs(Tvectors[i][j]) > error_tol) ritz[i] += Tvectors[i][j] * vec2;
    }
    for (size_t i = 0; i < ritz.size(); ++i) {
      *out = ritz[i];
      ++out;
      mult(matrix_, ritz[i], vec3);
      vec3 -= lambdas[i] * ritz[i];
      residuum.push_back(two_norm(vec3));
    }
    inf = Info<magnitude_type>(M1, M2, Ma, lambdas, residuum, status);
  }

  const std::vector<std::vector<magnitude_type>>& t_eigenvectors() { return Tvectors; }

  // Current pair (v_{n-1}, v_n) of Lanczos vectors.
  const vector_type& previous_vector() const { return startvector; }
  const vector_type& current_vector() const { return vec2; }
  unsigned int steps() const { return n; }

  // Archive the recurrence (src/ietl/lanczos.h:86-98): the T matrix, the
  // current pair of Lanczos vectors, the step count and the eigenvector
  // bookkeeping, in the reference's order.
  template <class Archive>
  void save(Archive& ar) const {
    static_cast<const super_type&>(*this).save(ar);
    ar << startvector << vec2 << n << M1 << M2 << Ma << Tvectors;
  }
  template <class Archive>
  void load(Archive& ar) {
    static_cast<super_type&>(*this).load(ar);
    ar >> startvector >> vec2 >> n >> M1 >> M2 >> Ma >> Tvectors;
  }

 private:
  template <class GEN>
  std::pair<magnitude_type, magnitude_type> make_first_step(GEN gen) {
    generate(startvector, gen);
    project(startvector, vecspace_);
    startvector /= two_norm(startvector);
    mult(matrix_, startvector, vec2);
    const magnitude_type a = real(dot(startvector, vec2));
    vec2 -= a * startvector;
    const magnitude_type b = two_norm(vec2);
    vec2 /= b;
    return std::make_pair(a, b);
  }

  std::pair<magnitude_type, magnitude_type> make_step(int j, vector_type& vec3) {
    const magnitude_type b = super_type::beta[j - 1];
    std::pair<magnitude_type, magnitude_type> ab = lanczos_step(matrix_, startvector, vec2, vec3, b);
    std::swap(vec2, startvector);
    std::swap(vec3, vec2);
    return ab;
  }

  template <class IT, class GEN>
  void generate_tmatrix(IT& iter, GEN gen) {
    if (super_type::alpha.size() == 0) {
      super_type::push_back(make_first_step(gen));
      n = 1;
    }
    generate_tmatrix(iter);
  }

  template <class IT>
  void generate_tmatrix(IT& iter) {
    vector_type vec3 = new_vector(vecspace_);
    for (unsigned int j = 0; j < n; ++j) ++iter;
    if (super_type::alpha.size() == 0)
      throw std::runtime_error(
          "T matrix error, size of T matrix is zero, more_eigenvalues() cannot be called.");
    while (!iter.finished(*this)) {
      std::pair<magnitude_type, magnitude_type> ab = make_step(n, vec3);
      if (n == super_type::alpha.size()) super_type::push_back(ab);
      ++iter;
      ++n;
    }
  }

  // m1 and m2 of each requested value: the first T lengths whose spectrum
  // has one, then two values within thold of it (lanczos.h:296-336)
  template <class IN>
  void find_m1m2(IN first, IN last) {
    const size_t cnt = (size_t)(last - first);
    M1.assign(cnt, 0);
    M2.assign(cnt, 0);
    const magnitude_type thold = super_type::thold_value();
    size_t m2counter = 0;
    unsigned int len = 1;
    while (m2counter < cnt && len < super_type::alpha.size()) {
      ++len;
      std::vector<magnitude_type> eval(super_type::alpha.begin(), super_type::alpha.begin() + len);
      if (tridiag::eig((int)len, eval.data(), super_type::beta.data(), nullptr) > 0)
        throw std::runtime_error("LAPACK error, stev function failed.");
      size_t i = 0;
      for (IN it = first; it != last; ++it, ++i) {
        if (M1[i] != 0 && M2[i] != 0) continue;
        const auto ub = std::lower_bound(eval.begin(), eval.end(), *it + thold);
        const auto lb = std::upper_bound(eval.begin(), eval.end(), *it - thold);
        if (M1[i] == 0 && ub - lb >= 1) M1[i] = (int)len;
        if (M2[i] == 0 && ub - lb >= 2) {
          M2[i] = (int)len;
          ++m2counter;
        }
      }
    }
  }

  const MATRIX& matrix_;
  const VS vecspace_;
  vector_type startvector;
  vector_type vec2;
  unsigned int n;
  std::vector<int> M1, M2, Ma;
  std::vector<std::vector<magnitude_type>> Tvectors;
};

}  // namespace ietl

#endif
