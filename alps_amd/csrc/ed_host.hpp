// ed_host.hpp -- host-side helpers shared by the C-ABI sources: error
// mapping to status codes, device buffers, model conversion and the
// T-matrix analysis of a solve.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <limits>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include <ietl/tmatrix.h>

#include "alps_ed.h"
#include "ed_layout.hpp"

namespace alps_ed {
namespace host {

// Per-thread error text of the last failed call and the stream calls
// enqueue on (alps_ed_last_error, alps_ed_set_stream).
inline std::string& last_error() {
  thread_local std::string s;
  return s;
}
inline hipStream_t& current_stream() {
  thread_local hipStream_t s = nullptr;
  return s;
}

struct DeviceError : std::runtime_error {
  explicit DeviceError(const std::string& s) : std::runtime_error(s) {}
};
struct MemoryError : std::runtime_error {
  explicit MemoryError(const std::string& s) : std::runtime_error(s) {}
};
struct StateError : std::runtime_error {
  explicit StateError(const std::string& s) : std::runtime_error(s) {}
};

inline void hip_check(hipError_t e, const char* what) {
  if (e == hipSuccess) return;
  if (e == hipErrorOutOfMemory) throw MemoryError(std::string(what) + ": " + hipGetErrorString(e));
  throw DeviceError(std::string(what) + ": " + hipGetErrorString(e));
}

template <class F>
int guarded(F&& f) {
  try {
    f();
    return ALPS_ED_OK;
  } catch (const std::invalid_argument& e) {
    last_error() = e.what();
    return ALPS_ED_ERR_ARGUMENT;
  } catch (const MemoryError& e) {
    last_error() = e.what();
    return ALPS_ED_ERR_MEMORY;
  } catch (const DeviceError& e) {
    last_error() = e.what();
    return ALPS_ED_ERR_DEVICE;
  } catch (const StateError& e) {
    last_error() = e.what();
    return ALPS_ED_ERR_STATE;
  } catch (const std::bad_alloc&) {
    last_error() = "host allocation failed";
    return ALPS_ED_ERR_MEMORY;
  } catch (const std::runtime_error& e) {
    last_error() = e.what();
    return ALPS_ED_ERR_SOLVER;
  } catch (const std::logic_error& e) {
    // a broken internal invariant of the host code (not a device fault)
    last_error() = std::string("internal error: ") + e.what();
    return ALPS_ED_ERR_SOLVER;
  } catch (const std::exception& e) {
    last_error() = e.what();
    return ALPS_ED_ERR_SOLVER;
  } catch (...) {
    last_error() = "unknown error";
    return ALPS_ED_ERR_DEVICE;
  }
}

inline void require(bool c, const char* msg) {
  if (!c) throw std::invalid_argument(msg);
}

// RAII device buffer
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  explicit DevBuf(size_t b) { alloc(b); }
  void alloc(size_t b) {
    release();
    if (b == 0) return;
    hip_check(hipMalloc(&p, b), "hipMalloc");
    bytes = b;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  ~DevBuf() { release(); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  template <class T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

// Device buffers of the kept Lanczos bases, recycled between solves of the
// process instead of freed: a basis chunk of an L=24 sector is 346 MB, and
// its hipMalloc would otherwise stall the host (and with it the queue of
// steps) once per 32 steps of every solve.  At most kMaxBytes stay cached;
// buffers are reused only at their exact size.
class BufferPool {
 public:
  static BufferPool& get() {
    static BufferPool p;
    return p;
  }
  void* acquire(size_t bytes) {
    {
      std::lock_guard<std::mutex> g(mu_);
      int dev = 0;
      (void)hipGetDevice(&dev);
      for (size_t i = 0; i < free_.size(); ++i)
        if (free_[i].bytes == bytes && free_[i].dev == dev) {
          void* p = free_[i].p;
          cached_ -= bytes;
          free_.erase(free_.begin() + i);
          return p;
        }
    }
    void* p = nullptr;
    hip_check(hipMalloc(&p, bytes), "hipMalloc");
    return p;
  }
  void release(void* p, size_t bytes) {
    if (!p) return;
    std::lock_guard<std::mutex> g(mu_);
    int dev = 0;
    (void)hipGetDevice(&dev);
    if (cached_ + bytes <= kMaxBytes) {
      free_.push_back({p, bytes, dev});
      cached_ += bytes;
    } else {
      (void)hipFree(p);
    }
  }

 private:
  static constexpr size_t kMaxBytes = size_t(4) << 30;
  struct Entry {
    void* p;
    size_t bytes;
    int dev;
  };
  std::mutex mu_;
  std::vector<Entry> free_;
  size_t cached_ = 0;
};

// A device buffer from the pool (returned to it when destroyed).
struct PooledBuf {
  void* p = nullptr;
  size_t bytes = 0;
  explicit PooledBuf(size_t b) : p(BufferPool::get().acquire(b)), bytes(b) {}
  ~PooledBuf() { BufferPool::get().release(p, bytes); }
  PooledBuf(const PooledBuf&) = delete;
  PooledBuf& operator=(const PooledBuf&) = delete;
};

// RAII mapped pinned host buffer: `h` for the host, `d` for kernels, which
// write it directly (small scalars read by the host after an event)
struct MappedBuf {
  void* h = nullptr;
  void* d = nullptr;
  MappedBuf() = default;
  void alloc(size_t b) {
    hip_check(hipHostMalloc(&h, b ? b : 1, hipHostMallocMapped), "hipHostMalloc");
    hip_check(hipHostGetDevicePointer(&d, h, 0), "hipHostGetDevicePointer");
  }
  ~MappedBuf() {
    if (h) (void)hipHostFree(h);
  }
  MappedBuf(const MappedBuf&) = delete;
  MappedBuf& operator=(const MappedBuf&) = delete;
};

// RAII HIP event without timing
struct Event {
  hipEvent_t e = nullptr;
  Event() { hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate"); }
  ~Event() {
    if (e) (void)hipEventDestroy(e);
  }
  Event(const Event&) = delete;
  Event& operator=(const Event&) = delete;
};

inline ModelSpec model_spec(const alps_ed_model* m) {
  require(m != nullptr, "model is null");
  require(m->n_sites >= 1 && m->n_sites <= kMaxSites, "n_sites must be in [1, 32]");
  require(m->n_bonds >= 0, "n_bonds must be >= 0");
  require(m->n_bonds == 0 || (m->bond_site1 && m->bond_site2), "bond site arrays are null");
  ModelSpec s;
  s.n_sites = m->n_sites;
  s.n_up = m->n_up;
  for (int b = 0; b < m->n_bonds; ++b) {
    require(m->bond_site1[b] >= 0 && m->bond_site1[b] < m->n_sites && m->bond_site2[b] >= 0 &&
                m->bond_site2[b] < m->n_sites && m->bond_site1[b] != m->bond_site2[b],
            "bond with invalid sites");
    s.bonds.push_back(make_bond(m->bond_site1[b], m->bond_site2[b], m->jz ? m->jz[b] : 0.0,
                                m->jxy ? m->jxy[b] : 0.0, m->k ? m->k[b] : 0.0));
  }
  for (int i = 0; i < m->n_sites; ++i) s.sites.push_back(make_site(m->h ? m->h[i] : 0.0, m->d ? m->d[i] : 0.0));
  return s;
}

// The lowest distinct non-ghost eigenvalues of T[0..m) with their error
// bounds and eigenvector positions.  With FP32-stored Lanczos vectors the
// recurrence produces near-copies of eigenvalues that differ by
// O(eps_f32 |T|), above the FP64 merge threshold; they are merged here at
// 64 eps_f32 |lambda|.  Unlike the FP64 rule (tmatrix.h:164-183, where an
// exact copy means convergence and the error is 0), FP32 near-copies also
// appear before convergence, so a merged group keeps the smallest error
// bound of its members and the position of its lowest member.
struct TValues {
  std::vector<double> ev, err;
  std::vector<int> index;
  std::vector<double> all;  // every eigenvalue of T, ascending (index refers to it)
};

inline TValues analyze_prefix(const std::vector<double>& a, const std::vector<double>& b, int m, bool fp32) {
  double amax = a[0], bmax = b[0];
  for (int i = 1; i < m; ++i) {
    amax = std::max(amax, a[i]);
    bmax = std::max(bmax, b[i]);
  }
  ietl::tmatrix_analysis an;
  an.run(a.data(), b.data(), m, amax, bmax, true);
  TValues t;
  t.all = std::move(an.eigval_all);
  const double dup = 64.0 * std::numeric_limits<float>::epsilon();
  for (size_t i = 0; i < an.eigval_noghost.size(); ++i) {
    const double lam = an.eigval_noghost[i];
    if (fp32 && !t.ev.empty() && lam - t.ev.back() <= dup * std::max(1.0, std::fabs(t.ev.back()))) {
      t.err.back() = std::min(t.err.back(), an.err_noghost[i]);
      continue;
    }
    t.ev.push_back(lam);
    t.err.push_back(an.err_noghost[i]);
    t.index.push_back(an.index_noghost[i]);
  }
  return t;
}

// Stopping tolerance (rtol = atol) of the nlowest rule: 100 eps, as
// lanczos_iteration_nlowest (iteration.h:110-113), for FP64 storage.  With
// FP32 storage the T-matrix error bounds stall near the FP32 noise level
// long before 100 eps; the solve stops at 1e-8 and reports FP64 Rayleigh
// quotients of the Ritz vectors, whose error is about the square of the
// residual (1e-14 and below on the tested chains and lattices).
inline double stop_tolerance(bool fp32) {
  return fp32 ? 1e-8 : 100. * std::numeric_limits<double>::epsilon();
}

// lanczos_iteration_nlowest::converged (iteration.h:116-130).
inline bool nlowest_converged(const TValues& t, int nev, bool fp32 = false) {
  const double tol = stop_tolerance(fp32);
  if ((int)t.ev.size() < nev) return false;
  for (int i = 0; i < nev; ++i)
    if (t.err[i] > std::max(tol, tol * std::fabs(t.ev[i]))) return false;
  return true;
}

// Cheap necessary-condition test for the nlowest stopping rule: the error
// bound |beta_{m-1} z_{m-1}| of the lowest eigenvalue of T[0..m), from
// Sturm-sequence bisection for the eigenvalue and the backward three-term
// recurrence for its eigenvector (stable exactly when the Ritz value has
// converged, i.e. when the last components are small).  O(m) per call.
// An optional bracket [hint_lo, hint_hi] (e.g. from the same test a few
// steps earlier: by interlacing the lowest value only moves down) shortens
// the bisection; it is used only if the Sturm counts confirm it.
inline double lowest_ritz_error(const double* a, const double* b, int m, double* lambda_out,
                                double hint_lo = 0.0, double hint_hi = -1.0) {
  double lo = a[0], hi = a[0];
  for (int i = 0; i < m; ++i) {
    const double r = (i > 0 ? std::fabs(b[i - 1]) : 0.0) + (i + 1 < m ? std::fabs(b[i]) : 0.0);
    lo = std::min(lo, a[i] - r);
    hi = std::max(hi, a[i] + r);
  }
  const double eps = std::numeric_limits<double>::epsilon();
  const double tiny = std::numeric_limits<double>::min();
  auto count_below = [&](double x) {  // eigenvalues of T[0..m) below x
    int c = 0;
    double q = a[0] - x;
    if (q < 0) ++c;
    for (int i = 1; i < m; ++i) {
      if (std::fabs(q) < tiny) q = q < 0 ? -tiny : tiny;
      q = a[i] - x - b[i - 1] * b[i - 1] / q;
      if (q < 0) ++c;
    }
    return c;
  };
  if (hint_lo < hint_hi && hint_lo >= lo && hint_hi <= hi && count_below(hint_lo) == 0 && count_below(hint_hi) >= 1) {
    lo = hint_lo;
    hi = hint_hi;
  }
  for (int it = 0; it < 200 && hi - lo > 2 * eps * std::max(std::fabs(lo), std::fabs(hi)); ++it) {
    const double mid = 0.5 * (lo + hi);
    if (mid <= lo || mid >= hi) break;
    if (count_below(mid) >= 1) hi = mid;
    else lo = mid;
  }
  const double lam = 0.5 * (lo + hi);
  if (lambda_out) *lambda_out = lam;
  // z_{m-1} = 1, then z_{i-1} = ((lam - a_i) z_i - b_i z_{i+1}) / b_{i-1}
  double z1 = 1.0, z2 = 0.0, nrm2 = 1.0, zlast = 1.0;
  for (int i = m - 1; i >= 1; --i) {
    const double zi = ((lam - a[i]) * z1 - (i + 1 < m ? b[i] * z2 : 0.0)) / b[i - 1];
    z2 = z1;
    z1 = zi;
    nrm2 += zi * zi;
    if (nrm2 > 1e200) {  // rescale: only the ratio zlast / |z| matters
      const double s = 1e-100;
      z1 *= s;
      z2 *= s;
      zlast *= s;
      nrm2 *= s * s;
    }
  }
  return std::fabs(b[m - 1]) * std::fabs(zlast) / std::sqrt(nrm2);
}

// Number of eigenvalues of T[0..m) below x (Sturm sequence of the LDL^T
// pivots).
inline int sturm_count_below(const double* a, const double* b, int m, double x) {
  const double tiny = std::numeric_limits<double>::min();
  int c = 0;
  double q = a[0] - x;
  if (q < 0) ++c;
  for (int i = 1; i < m; ++i) {
    if (std::fabs(q) < tiny) q = q < 0 ? -tiny : tiny;
    q = a[i] - x - b[i - 1] * b[i - 1] / q;
    if (q < 0) ++c;
  }
  return c;
}

// A necessary condition for nlowest_converged(analyze_prefix(a, b, m, fp32),
// nev, fp32), in O(m): the lowest eigenvalue of T[0..m) has an error bound
// within 4 times the stopping tolerance (the slack covers the difference
// between this bound's backward recurrence and the full analysis'
// eigenvectors), or another eigenvalue of T lies within
// 1e-5 max(|T|, 1) of it.  In the second case the full analysis may merge
// the two (the FP64 rule's thold, tmatrix.h:149-150, is about 1e-10 |T|; the
// FP32 near-copy window 64 eps_f32 max(|lambda|, 1) is below 1e-5 max(|T|,
// 1)) and give the group
// a smaller bound, so the full analysis has to decide.  When the test fails
// the stopping rule cannot hold, because the lowest distinct value of the
// analysis is then this eigenvalue with this bound.
inline bool may_be_converged(const double* a, const double* b, int m, bool fp32, double* lambda_out = nullptr,
                             double hint_lo = 0.0, double hint_hi = -1.0) {
  double lam = 0.0;
  const double err = lowest_ritz_error(a, b, m, &lam, hint_lo, hint_hi);
  if (lambda_out) *lambda_out = lam;
  const double tol = stop_tolerance(fp32);
  if (!(err > 4.0 * std::max(tol, tol * std::fabs(lam)))) return true;  // also NaN: the full analysis reports it
  double tn = 0.0;
  for (int i = 0; i < m; ++i)
    tn = std::max(tn, std::fabs(a[i]) + (i > 0 ? std::fabs(b[i - 1]) : 0.0) + (i + 1 < m ? std::fabs(b[i]) : 0.0));
  return sturm_count_below(a, b, m, lam + 1e-5 * std::max({tn, std::fabs(lam), 1.0})) >= 2;
}

// Eigenvector of T[0..m) for the (already accurate) eigenvalue lambda by
// inverse iteration: Gaussian elimination with partial pivoting on the
// tridiagonal T - lambda I (zero pivots replaced by eps |T|), three solves
// from a fixed start, normalized.  z has m entries.
inline void tridiag_eigenvector(const double* a, const double* b, int m, double lambda, double* z) {
  const double eps = std::numeric_limits<double>::epsilon();
  double tnorm = 0.0;
  for (int i = 0; i < m; ++i)
    tnorm = std::max(tnorm, std::fabs(a[i]) + (i > 0 ? std::fabs(b[i - 1]) : 0.0) + (i + 1 < m ? std::fabs(b[i]) : 0.0));
  const double small = eps * std::max(tnorm, std::numeric_limits<double>::min());
  // LU of T - lambda I with row interchanges (the dgttrf scheme): U has the
  // diagonals dd, du, du2; dl holds the multipliers.
  std::vector<double> dd(m), dl(m, 0.0), du(m, 0.0), du2(m, 0.0);
  std::vector<char> swp(m, 0);
  for (int i = 0; i < m; ++i) {
    dd[i] = a[i] - lambda;
    if (i + 1 < m) dl[i] = du[i] = b[i];
  }
  auto guard = [&](double v) { return std::fabs(v) < small ? (v < 0 ? -small : small) : v; };
  for (int i = 0; i + 1 < m; ++i) {
    if (std::fabs(dd[i]) >= std::fabs(dl[i])) {
      dd[i] = guard(dd[i]);
      const double f = dl[i] / dd[i];
      dl[i] = f;
      dd[i + 1] -= f * du[i];
    } else {
      const double f = dd[i] / dl[i];
      dd[i] = dl[i];
      dl[i] = f;
      const double t = du[i];
      du[i] = dd[i + 1];
      dd[i + 1] = t - f * dd[i + 1];
      if (i + 2 < m) {
        du2[i] = du[i + 1];
        du[i + 1] = -f * du[i + 1];
      }
      swp[i] = 1;
    }
  }
  dd[m - 1] = guard(dd[m - 1]);
  std::vector<double> x(m, 1.0);
  for (int it = 0; it < 3; ++it) {
    // forward: the row interchanges and multipliers
    for (int i = 0; i + 1 < m; ++i) {
      if (swp[i]) std::swap(x[i], x[i + 1]);
      x[i + 1] -= dl[i] * x[i];
    }
    // back substitution
    for (int i = m - 1; i >= 0; --i) {
      double v = x[i];
      if (i + 1 < m) v -= du[i] * x[i + 1];
      if (i + 2 < m) v -= du2[i] * x[i + 2];
      x[i] = v / dd[i];
    }
    double nrm = 0.0;
    for (double v : x) nrm += v * v;
    nrm = std::sqrt(nrm);
    for (double& v : x) v /= nrm;
  }
  for (int i = 0; i < m; ++i) z[i] = x[i];
}

}  // namespace host
}  // namespace alps_ed
