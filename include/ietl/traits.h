// ietl/traits.h -- vector-space traits of the ietl concept
// (same names and meaning as src/ietl/traits.h:23-40 of the reference).
#ifndef ALPS_ED_IETL_TRAITS_H
#define ALPS_ED_IETL_TRAITS_H

#include <complex>

namespace ietl {

template <class T>
struct number_traits {
  typedef T magnitude_type;
};

template <class T>
struct number_traits<std::complex<T> > {
  typedef T magnitude_type;
};

template <class VS>
struct vectorspace_traits {
  typedef typename VS::vector_type vector_type;
  typedef typename VS::size_type size_type;
  typedef typename VS::scalar_type scalar_type;
  typedef typename number_traits<scalar_type>::magnitude_type magnitude_type;
};

inline double real(double x) { return x; }
inline float real(float x) { return x; }
template <class T>
inline T real(const std::complex<T>& x) { return x.real(); }

}  // namespace ietl

#endif
