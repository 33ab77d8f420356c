// alps_ed/ietl_gpu.hpp -- ietl vector space and operator on the MI355X engine.
//
// Lets ietl::lanczos<MATRIX, VS> (src/ietl/lanczos.h:64-119 of the reference,
// or include/ietl/lanczos.h here) run with the Hamiltonian and the Lanczos
// vectors in HBM:
//     alps_ed::hamiltonian H(model);                  // MATRIX
//     alps_ed::gpu_vectorspace vs(H.dimension());     // VS (FP32 storage)
//     ietl::lanczos<alps_ed::hamiltonian, alps_ed::gpu_vectorspace> lz(H, vs);
//     ietl::lanczos_iteration_nlowest<double> it(max_iter, nev);
//     lz.calculate_eigenvalues(it, alps_ed::uniform_start{seed});
// The concept pieces follow src/ietl/vectorspace.h:23-46,101-109,
// src/ietl/traits.h:23-40 and src/ietl/interface/ublas.h:27-63:
//   vectorspace_traits typedefs; new_vector, vec_dimension, project;
//   generate, dot, two_norm, copy, mult; v *= s, v /= s, v -= s*w, v += s*w,
//   v = s*w; std::swap.
// Vectors are handles (as ietl::vector_wrapper, vectorspace.h:51-99): a copy
// shares the device buffer, swap is O(1).  scalar_type is double: dot
// products and norms accumulate in FP64 whatever the storage type.
// Errors of the engine are thrown as std::invalid_argument (bad arguments)
// or std::runtime_error (everything else), the exceptions ietl itself uses.
//
// Include this header before <ietl/lanczos.h> when the reference's ietl is
// used, so that its qualified ietl:: calls see the overloads below.
#ifndef ALPS_ED_IETL_GPU_HPP
#define ALPS_ED_IETL_GPU_HPP

#include <alps_ed.h>

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace alps_ed {

inline void check(int status) {
  if (status == ALPS_ED_OK) return;
  const std::string msg = alps_ed_last_error();
  if (status == ALPS_ED_ERR_ARGUMENT) throw std::invalid_argument(msg);
  throw std::runtime_error(msg);
}

// Device Hamiltonian of one fixed-Sz sector (the MATRIX of ietl::lanczos).
class hamiltonian {
 public:
  explicit hamiltonian(const alps_ed_model& m) {
    alps_ed_operator h = nullptr;
    check(alps_ed_operator_create(&m, &h));
    op_.reset(h, [](alps_ed_operator p) { alps_ed_operator_destroy(p); });
    check(alps_ed_operator_dimension(h, &dim_));
  }
  uint64_t dimension() const { return dim_; }
  alps_ed_operator handle() const { return op_.get(); }

 private:
  std::shared_ptr<alps_ed_operator_s> op_;
  uint64_t dim_ = 0;
};

class gpu_vector;

struct scaled_gpu_vector {
  const gpu_vector& v;
  double s;
};

// Device vector with handle semantics.
class gpu_vector {
 public:
  typedef double value_type;
  typedef uint64_t size_type;

  gpu_vector() = default;
  gpu_vector(uint64_t n, alps_ed_dtype t) {
    alps_ed_vector v = nullptr;
    check(alps_ed_vector_create(n, t, &v));
    v_.reset(v, [](alps_ed_vector p) { alps_ed_vector_destroy(p); });
    n_ = n;
    t_ = t;
  }
  uint64_t size() const { return n_; }
  alps_ed_dtype dtype() const { return t_; }
  alps_ed_vector handle() const { return v_.get(); }

  gpu_vector& operator*=(double a) {
    check(alps_ed_vector_scale(handle(), a));
    return *this;
  }
  gpu_vector& operator/=(double a) { return *this *= 1.0 / a; }
  gpu_vector& operator+=(const scaled_gpu_vector& x) {
    check(alps_ed_vector_axpby(handle(), x.s, x.v.handle(), 1.0));
    return *this;
  }
  gpu_vector& operator-=(const scaled_gpu_vector& x) {
    check(alps_ed_vector_axpby(handle(), -x.s, x.v.handle(), 1.0));
    return *this;
  }
  gpu_vector& operator=(const scaled_gpu_vector& x) {
    check(alps_ed_vector_axpby(handle(), x.s, x.v.handle(), 0.0));
    return *this;
  }
  std::vector<double> to_host() const {
    std::vector<double> h(n_);
    check(alps_ed_vector_download(handle(), h.data(), n_));
    return h;
  }
  void from_host(const std::vector<double>& h) { check(alps_ed_vector_upload(handle(), h.data(), h.size())); }

 private:
  std::shared_ptr<alps_ed_vector_s> v_;
  uint64_t n_ = 0;
  alps_ed_dtype t_ = ALPS_ED_F32;
};

inline scaled_gpu_vector operator*(double s, const gpu_vector& v) { return {v, s}; }
inline scaled_gpu_vector operator*(const gpu_vector& v, double s) { return {v, s}; }
inline scaled_gpu_vector operator/(const gpu_vector& v, double s) { return {v, 1.0 / s}; }

// The VS of ietl::lanczos: vectors of one dimension in FP32 or FP64 storage.
class gpu_vectorspace {
 public:
  typedef gpu_vector vector_type;
  typedef double scalar_type;
  typedef uint64_t size_type;

  explicit gpu_vectorspace(uint64_t n, alps_ed_dtype t = ALPS_ED_F32) : n_(n), t_(t) {}
  size_type vec_dimension() const { return n_; }
  vector_type new_vector() const { return gpu_vector(n_, t_); }
  void project(vector_type&) const {}
  alps_ed_dtype dtype() const { return t_; }

 private:
  uint64_t n_;
  alps_ed_dtype t_;
};

// Start-vector generator filled on the device: uniform on [lo, hi) from a
// counter-based splitmix64 stream (alps_ed_vector_fill_random).  Like the
// reference's generator passed by value, it regenerates the same vector on
// every call.
struct uniform_start {
  uint64_t seed = 0;
  double lo = -1.0, hi = 1.0;
};

// Free functions of the ietl concept, found by argument-dependent lookup.
inline void generate(gpu_vector& v, const uniform_start& g) {
  check(alps_ed_vector_fill_random(v.handle(), g.seed, g.lo, g.hi));
}
// Any other generator runs on the host (std::generate semantics, with the
// generator copied so each call restarts it) and the values are uploaded.
template <class Gen>
void generate(gpu_vector& v, const Gen& gen) {
  Gen g = gen;
  std::vector<double> h(v.size());
  for (auto& x : h) x = g();
  v.from_host(h);
}
inline double dot(const gpu_vector& x, const gpu_vector& y) {
  double r = 0.0;
  check(alps_ed_vector_dot(x.handle(), y.handle(), &r));
  return r;
}
inline double two_norm(const gpu_vector& x) {
  double r = 0.0;
  check(alps_ed_vector_nrm2(x.handle(), &r));
  return r;
}
inline void copy(const gpu_vector& x, gpu_vector& y) { check(alps_ed_vector_copy(y.handle(), x.handle())); }
inline void mult(const hamiltonian& h, const gpu_vector& x, gpu_vector& y) {
  check(alps_ed_matvec(h.handle(), x.handle(), y.handle()));
}
inline void clear(gpu_vector& v) { check(alps_ed_vector_fill(v.handle(), 0.0)); }
inline gpu_vector new_vector(const gpu_vectorspace& vs) { return vs.new_vector(); }
inline uint64_t vec_dimension(const gpu_vectorspace& vs) { return vs.vec_dimension(); }
inline void project(gpu_vector&, const gpu_vectorspace&) {}

// One make_step as a fused device call (used by include/ietl/lanczos.h):
// w = H v_cur - alpha v_cur - beta_prev v_prev, normalized; returns (alpha, beta).
inline std::pair<double, double> lanczos_step(const hamiltonian& h, gpu_vector& v_prev, gpu_vector& v_cur,
                                              gpu_vector& w, double beta_prev) {
  double a = 0.0, b = 0.0;
  check(alps_ed_lanczos_step_vectors(h.handle(), v_prev.handle(), v_cur.handle(), w.handle(), beta_prev, &a, &b));
  return std::make_pair(a, b);
}

}  // namespace alps_ed

// Qualified ietl:: spellings of the same functions, for templates that call
// ietl::mult, ietl::dot, ietl::two_norm, ietl::generate explicitly.
namespace ietl {
using alps_ed::copy;
using alps_ed::dot;
using alps_ed::generate;
using alps_ed::mult;
using alps_ed::two_norm;
}  // namespace ietl

#endif
