// ietl/vectorspace.h -- the ietl vector-space concept for host vectors
// (src/ietl/vectorspace.h:23-46,101-109 and src/ietl/interface/ublas.h:27-63
// of the reference, here over std::vector so that no Boost is needed).
// Device vector spaces live in alps_ed/ietl_gpu.hpp.
#ifndef ALPS_ED_IETL_VECTORSPACE_H
#define ALPS_ED_IETL_VECTORSPACE_H

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <vector>

#include <ietl/traits.h>

namespace ietl {

template <class V>
class vectorspace {
 public:
  typedef V vector_type;
  typedef typename V::value_type scalar_type;
  typedef typename V::size_type size_type;

  explicit vectorspace(size_type n) : n_(n) {}
  size_type vec_dimension() const { return n_; }
  vector_type new_vector() const { return vector_type(n_, scalar_type()); }
  void project(vector_type&) const {}

 private:
  size_type n_;
};

template <class VS>
typename vectorspace_traits<VS>::vector_type new_vector(const VS& vs) {
  return vs.new_vector();
}

template <class VS>
typename vectorspace_traits<VS>::size_type vec_dimension(const VS& vs) {
  return vs.vec_dimension();
}

template <class VS>
void project(typename vectorspace_traits<VS>::vector_type& v, const VS& vs) {
  vs.project(v);
}

// std::vector<T> bindings: generate (std::generate semantics), dot, two_norm,
// copy, and the expressions the recurrence uses.
template <class T, class Gen>
inline void generate(std::vector<T>& c, const Gen& gen) {
  Gen g = gen;  // the generator is taken by value, so every call restarts it
  std::generate(c.begin(), c.end(), g);
}

template <class T>
inline T dot(const std::vector<T>& x, const std::vector<T>& y) {
  T t = T();
  for (std::size_t i = 0; i < x.size(); ++i) t += x[i] * y[i];
  return t;
}

template <class T>
inline T two_norm(const std::vector<T>& x) {
  T t = T();
  for (std::size_t i = 0; i < x.size(); ++i) t += x[i] * x[i];
  return std::sqrt(t);
}

template <class T>
inline void copy(const std::vector<T>& x, std::vector<T>& y) {
  y = x;
}

template <class T, class S>
struct scaled_host_vector {
  const std::vector<T>& v;
  S s;
};

template <class T, class S>
inline scaled_host_vector<T, S> operator*(S s, const std::vector<T>& v) {
  return {v, s};
}

template <class T, class S>
inline std::vector<T>& operator-=(std::vector<T>& y, const scaled_host_vector<T, S>& x) {
  for (std::size_t i = 0; i < y.size(); ++i) y[i] -= x.s * x.v[i];
  return y;
}

template <class T, class S>
inline std::vector<T>& operator+=(std::vector<T>& y, const scaled_host_vector<T, S>& x) {
  for (std::size_t i = 0; i < y.size(); ++i) y[i] += x.s * x.v[i];
  return y;
}

template <class T, class S>
inline std::vector<T>& operator/=(std::vector<T>& y, S s) {
  for (auto& v : y) v /= s;
  return y;
}

template <class T, class S>
inline std::vector<T>& operator*=(std::vector<T>& y, S s) {
  for (auto& v : y) v *= s;
  return y;
}

}  // namespace ietl

#endif
