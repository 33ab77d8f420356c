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

#include "alps_ed.h"
#include "ed_layout.hpp"
#include "ed_tridiag.hpp"

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
}  // namespace host
}  // namespace alps_ed
