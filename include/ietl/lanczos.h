// ietl/lanczos.h -- ietl::lanczos<MATRIX, VS> with the reference's interface
// and recurrence (src/ietl/lanczos.h:64-129,341-406):
//   make_first_step   v0 = gen; project; v0 /= |v0|; w = H v0;
//                     a = Re<v0,w>; w -= a v0; b = |w|; v1 = w/b
//   make_step(j)      w = H v_j; a = Re<v_j,w>; w -= a v_j;
//                     w -= beta[j-1] v_{j-1}; b = |w|; w /= b; rotate
//   generate_tmatrix  first step when T is empty, then make_step while
//                     !iter.finished(*this)
// No reorthogonalization, exactly as the reference.  The only addition is a
// customization point: the step body is `lanczos_step(matrix, v_prev, v_cur,
// w, beta_prev)`, found by argument-dependent lookup, so that a device
// vector space can run the step as fused kernels (alps_ed/ietl_gpu.hpp).
// The generic version below performs the reference's operation sequence.
#ifndef ALPS_ED_IETL_LANCZOS_H
#define ALPS_ED_IETL_LANCZOS_H

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

  // Current pair (v_{n-1}, v_n) of Lanczos vectors.
  const vector_type& previous_vector() const { return startvector; }
  const vector_type& current_vector() const { return vec2; }
  unsigned int steps() const { return n; }

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

  const MATRIX& matrix_;
  const VS vecspace_;
  vector_type startvector;
  vector_type vec2;
  unsigned int n;
};

}  // namespace ietl

#endif
