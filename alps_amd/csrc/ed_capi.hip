// ed_capi.hip -- implementation of the C ABI declared in include/alps_ed.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include <ietl/tmatrix.h>

#include "alps_ed.h"
#include "ed_kernels.hpp"

using namespace alps_ed;

namespace {

thread_local std::string g_last_error;
thread_local hipStream_t g_stream = nullptr;

struct DeviceError : std::runtime_error {
  explicit DeviceError(const std::string& s) : std::runtime_error(s) {}
};
struct MemoryError : std::runtime_error {
  explicit MemoryError(const std::string& s) : std::runtime_error(s) {}
};
struct StateError : std::runtime_error {
  explicit StateError(const std::string& s) : std::runtime_error(s) {}
};

void hip_check(hipError_t e, const char* what) {
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
    g_last_error = e.what();
    return ALPS_ED_ERR_ARGUMENT;
  } catch (const MemoryError& e) {
    g_last_error = e.what();
    return ALPS_ED_ERR_MEMORY;
  } catch (const DeviceError& e) {
    g_last_error = e.what();
    return ALPS_ED_ERR_DEVICE;
  } catch (const StateError& e) {
    g_last_error = e.what();
    return ALPS_ED_ERR_STATE;
  } catch (const std::bad_alloc&) {
    g_last_error = "host allocation failed";
    return ALPS_ED_ERR_MEMORY;
  } catch (const std::runtime_error& e) {
    g_last_error = e.what();
    return ALPS_ED_ERR_SOLVER;
  } catch (...) {
    g_last_error = "unknown error";
    return ALPS_ED_ERR_DEVICE;
  }
}

void require(bool c, const char* msg) {
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

template <class T>
void upload(DevBuf& b, const std::vector<T>& v) {
  b.alloc(std::max<size_t>(v.size(), 1) * sizeof(T));
  if (!v.empty()) hip_check(hipMemcpy(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice), "upload");
}

// Scratch for grid reductions: partial sums, two counters, result slots.
struct Reducer {
  DevBuf partials, counters, results;
  size_t n_partials = 0;
  void ensure(size_t nblocks) {
    if (nblocks <= n_partials) return;
    partials.alloc(nblocks * sizeof(double));
    n_partials = nblocks;
    if (!counters.p) {
      counters.alloc(4 * sizeof(unsigned));
      hip_check(hipMemset(counters.p, 0, 4 * sizeof(unsigned)), "hipMemset");
      hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
      results.alloc(8 * sizeof(double));
    }
  }
  GridReduce make(int counter, double* out, double* out2, int mode) {
    GridReduce r;
    r.partials = partials.as<double>();
    r.counter = counters.as<unsigned>() + counter;
    r.out = out;
    r.out2 = out2;
    r.mode = mode;
    return r;
  }
  double* result(int i) { return results.as<double>() + i; }
};

Reducer& thread_reducer() {
  thread_local Reducer r;
  r.ensure(4096);
  return r;
}

}  // namespace

// ---------------------------------------------------------------------------
// object definitions
// ---------------------------------------------------------------------------

struct alps_ed_operator_s {
  SectorLayout layout;
  DevLayout dev{};
  DevBuf lo_rank, lo_state, lo_class_off, hi_base, seg_hi, seg_start, tiles, diag_lo, diag_hi, ll, hh,
      xb, site_dn, site_up, bond_p1, bond_p2, bond_par, bond_anti, bond_t;
  uint64_t bytes = 0;
};

struct alps_ed_vector_s {
  uint64_t n = 0;
  alps_ed_dtype dtype = ALPS_ED_F64;
  DevBuf buf;
};

struct alps_ed_lanczos_s {
  alps_ed_operator op = nullptr;
  alps_ed_dtype dtype = ALPS_ED_F32;
  int max_steps = 0;
  int steps = 0;  // alpha entries produced
  DevBuf u[2], w;
  DevBuf alpha, beta, scale;  // FP64 on device
  Reducer red;
  hipStream_t stream = nullptr;
};

namespace {

ModelSpec model_spec(const alps_ed_model* m) {
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

void operator_upload(alps_ed_operator op) {
  SectorLayout& L = op->layout;
  upload(op->lo_rank, L.lo_rank);
  upload(op->lo_state, L.lo_state);
  upload(op->lo_class_off, L.lo_class_off);
  upload(op->hi_base, L.hi_base);
  upload(op->seg_hi, L.seg_hi);
  upload(op->seg_start, L.seg_start);
  upload(op->tiles, L.tiles);
  upload(op->diag_lo, L.diag_lo);
  upload(op->diag_hi, L.diag_hi);
  upload(op->ll, L.ll);
  upload(op->hh, L.hh);
  upload(op->xb, L.xb);
  upload(op->site_dn, L.site_dn);
  upload(op->site_up, L.site_up);
  upload(op->bond_p1, L.bond_p1);
  upload(op->bond_p2, L.bond_p2);
  upload(op->bond_par, L.bond_par);
  upload(op->bond_anti, L.bond_anti);
  upload(op->bond_t, L.bond_t);
  DevLayout& d = op->dev;
  d.N = L.N;
  d.k = L.k;
  d.B = L.B;
  d.NH = L.NH;
  d.D = (uint32_t)L.D;
  d.n_seg = (uint32_t)L.seg_hi.size();
  d.n_tiles = (uint32_t)(L.tiles.size() / 2);
  d.n_ll = (int)L.ll.size();
  d.n_hh = (int)L.hh.size();
  d.n_xb = (int)L.xb.size();
  d.n_bonds = (int)L.bond_p1.size();
  d.lo_rank = op->lo_rank.as<uint32_t>();
  d.lo_state = op->lo_state.as<uint32_t>();
  d.lo_class_off = op->lo_class_off.as<uint32_t>();
  d.hi_base = op->hi_base.as<uint32_t>();
  d.seg_hi = op->seg_hi.as<uint32_t>();
  d.seg_start = op->seg_start.as<uint32_t>();
  d.tiles = op->tiles.as<uint32_t>();
  d.diag_lo = op->diag_lo.as<double>();
  d.diag_hi = op->diag_hi.as<double>();
  d.ll = op->ll.as<FlipBond>();
  d.hh = op->hh.as<FlipBond>();
  d.xb = op->xb.as<CrossBond>();
  d.site_dn = op->site_dn.as<double>();
  d.site_up = op->site_up.as<double>();
  d.bond_p1 = op->bond_p1.as<uint32_t>();
  d.bond_p2 = op->bond_p2.as<uint32_t>();
  d.bond_par = op->bond_par.as<double>();
  d.bond_anti = op->bond_anti.as<double>();
  d.bond_t = op->bond_t.as<double>();
  op->bytes = 0;
  for (DevBuf* b : {&op->lo_rank, &op->lo_state, &op->lo_class_off, &op->hi_base, &op->seg_hi,
                    &op->seg_start, &op->tiles, &op->diag_lo, &op->diag_hi, &op->ll, &op->hh, &op->xb,
                    &op->site_dn, &op->site_up, &op->bond_p1, &op->bond_p2, &op->bond_par,
                    &op->bond_anti, &op->bond_t})
    op->bytes += b->bytes;
}

size_t dtype_size(alps_ed_dtype t) { return t == ALPS_ED_F32 ? 4 : 8; }

void check_vec(alps_ed_vector v) { require(v != nullptr, "vector is null"); }

// dispatch on the storage type of a vector
template <class F>
void with_dtype(alps_ed_dtype t, F&& f) {
  if (t == ALPS_ED_F32) f(float());
  else f(double());
}

template <class V>
double dot_sync(uint64_t n, const V* x, const V* y) {
  Reducer& r = thread_reducer();
  hip_check(launch_dot<V>(n, x, y, r.make(0, r.result(0), nullptr, 0), g_stream), "dot");
  double h = 0.0;
  hip_check(hipMemcpyAsync(&h, r.result(0), sizeof(double), hipMemcpyDeviceToHost, g_stream), "copy");
  hip_check(hipStreamSynchronize(g_stream), "sync");
  return h;
}

// One Lanczos step of the device-resident recurrence (see alps_ed_lanczos_*).
template <class V>
void lanczos_step_dev(alps_ed_lanczos lz, int j) {
  const DevLayout& L = lz->op->dev;
  V* cur = lz->u[j & 1].as<V>();
  V* prev = lz->u[(j + 1) & 1].as<V>();
  V* w = lz->w.as<V>();
  double* alpha = lz->alpha.as<double>();
  double* beta = lz->beta.as<double>();
  double* scale = lz->scale.as<double>();
  hip_check(launch_hv<V>(L, cur, w, scale + j, lz->red.make(0, alpha + j, nullptr, 0), lz->stream), "hv");
  hip_check(launch_update<V>(L.D, w, cur, prev, prev, alpha + j, scale + j, j > 0 ? beta + j - 1 : nullptr,
                             j > 0 ? scale + j - 1 : nullptr,
                             lz->red.make(1, beta + j, scale + j + 1, 1), lz->stream),
            "update");
}

}  // namespace

// ---------------------------------------------------------------------------
// library
// ---------------------------------------------------------------------------
extern "C" {

int alps_ed_version(void) { return 100; }

const char* alps_ed_last_error(void) { return g_last_error.c_str(); }

int alps_ed_device_info(char* name, size_t name_len, uint64_t* hbm_bytes, int* n_cu) {
  return guarded([&] {
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    hipDeviceProp_t p;
    hip_check(hipGetDeviceProperties(&p, dev), "hipGetDeviceProperties");
    if (name && name_len) {
      std::string s = std::string(p.name) + " (" + p.gcnArchName + ")";
      std::strncpy(name, s.c_str(), name_len - 1);
      name[name_len - 1] = 0;
    }
    if (hbm_bytes) *hbm_bytes = p.totalGlobalMem;
    if (n_cu) *n_cu = p.multiProcessorCount;
  });
}

int alps_ed_set_stream(void* s) {
  g_stream = static_cast<hipStream_t>(s);
  return ALPS_ED_OK;
}

int alps_ed_synchronize(void) {
  return guarded([&] { hip_check(hipStreamSynchronize(g_stream), "hipStreamSynchronize"); });
}

// ---------------------------------------------------------------------------
// operator
// ---------------------------------------------------------------------------
int alps_ed_operator_create(const alps_ed_model* model, alps_ed_operator* out) {
  return guarded([&] {
    require(out != nullptr, "output handle is null");
    ModelSpec spec = model_spec(model);
    std::unique_ptr<alps_ed_operator_s> op(new alps_ed_operator_s);
    op->layout = build_layout(spec);
    operator_upload(op.get());
    *out = op.release();
  });
}

int alps_ed_operator_destroy(alps_ed_operator op) {
  delete op;
  return ALPS_ED_OK;
}

int alps_ed_operator_dimension(alps_ed_operator op, uint64_t* dim) {
  return guarded([&] {
    require(op && dim, "null argument");
    *dim = op->layout.D;
  });
}

int alps_ed_operator_bytes(alps_ed_operator op, uint64_t* bytes) {
  return guarded([&] {
    require(op && bytes, "null argument");
    *bytes = op->bytes;
  });
}

int alps_ed_operator_states(alps_ed_operator op, uint64_t first, uint64_t count, uint32_t* states_out) {
  return guarded([&] {
    require(op && (states_out || count == 0), "null argument");
    require(first + count <= op->layout.D, "state range beyond the dimension");
    if (count == 0) return;
    DevBuf d(count * sizeof(uint32_t));
    hip_check(launch_unrank(op->dev, first, count, d.as<uint32_t>(), g_stream), "unrank");
    hip_check(hipMemcpyAsync(states_out, d.p, count * sizeof(uint32_t), hipMemcpyDeviceToHost, g_stream), "copy");
    hip_check(hipStreamSynchronize(g_stream), "sync");
  });
}

int alps_ed_operator_index(alps_ed_operator op, const uint32_t* states, uint64_t count, uint64_t* index_out) {
  return guarded([&] {
    require(op && ((states && index_out) || count == 0), "null argument");
    if (count == 0) return;
    DevBuf ds(count * sizeof(uint32_t)), di(count * sizeof(uint64_t));
    hip_check(hipMemcpyAsync(ds.p, states, count * sizeof(uint32_t), hipMemcpyHostToDevice, g_stream), "copy");
    hip_check(launch_rank(op->dev, ds.as<uint32_t>(), count, di.as<uint64_t>(), g_stream), "rank");
    hip_check(hipMemcpyAsync(index_out, di.p, count * sizeof(uint64_t), hipMemcpyDeviceToHost, g_stream), "copy");
    hip_check(hipStreamSynchronize(g_stream), "sync");
  });
}

int alps_ed_operator_elements(alps_ed_operator op, uint64_t row0, uint64_t nrows, int max_per_row,
                              uint32_t* cols, double* vals, int* counts) {
  return guarded([&] {
    require(op && ((cols && vals && counts) || nrows == 0), "null argument");
    require(row0 + nrows <= op->layout.D, "row range beyond the dimension");
    require(max_per_row >= (int)op->layout.bond_p1.size() + 1, "max_per_row must be >= n_bonds + 1");
    if (nrows == 0) return;
    const uint64_t ne = nrows * (uint64_t)max_per_row;
    DevBuf dc(ne * sizeof(uint32_t)), dv(ne * sizeof(double)), dn(nrows * sizeof(int));
    hip_check(launch_elements(op->dev, row0, nrows, max_per_row, dc.as<uint32_t>(), dv.as<double>(),
                              dn.as<int>(), g_stream),
              "elements");
    hip_check(hipMemcpyAsync(cols, dc.p, ne * sizeof(uint32_t), hipMemcpyDeviceToHost, g_stream), "copy");
    hip_check(hipMemcpyAsync(vals, dv.p, ne * sizeof(double), hipMemcpyDeviceToHost, g_stream), "copy");
    hip_check(hipMemcpyAsync(counts, dn.p, nrows * sizeof(int), hipMemcpyDeviceToHost, g_stream), "copy");
    hip_check(hipStreamSynchronize(g_stream), "sync");
  });
}

// ---------------------------------------------------------------------------
// vectors
// ---------------------------------------------------------------------------
int alps_ed_vector_create(uint64_t n, alps_ed_dtype dtype, alps_ed_vector* v) {
  return guarded([&] {
    require(v != nullptr, "output handle is null");
    require(dtype == ALPS_ED_F32 || dtype == ALPS_ED_F64, "unknown dtype");
    std::unique_ptr<alps_ed_vector_s> p(new alps_ed_vector_s);
    p->n = n;
    p->dtype = dtype;
    p->buf.alloc(std::max<uint64_t>(n, 1) * dtype_size(dtype));
    with_dtype(dtype, [&](auto z) {
      using V = decltype(z);
      hip_check(launch_fill<V>(n, p->buf.as<V>(), 0.0, g_stream), "fill");
    });
    *v = p.release();
  });
}

int alps_ed_vector_destroy(alps_ed_vector v) {
  delete v;
  return ALPS_ED_OK;
}

int alps_ed_vector_size(alps_ed_vector v, uint64_t* n) {
  return guarded([&] {
    check_vec(v);
    require(n != nullptr, "null argument");
    *n = v->n;
  });
}

int alps_ed_vector_dtype(alps_ed_vector v, alps_ed_dtype* t) {
  return guarded([&] {
    check_vec(v);
    require(t != nullptr, "null argument");
    *t = v->dtype;
  });
}

int alps_ed_vector_data(alps_ed_vector v, void** p) {
  return guarded([&] {
    check_vec(v);
    require(p != nullptr, "null argument");
    *p = v->buf.p;
  });
}

int alps_ed_vector_upload(alps_ed_vector v, const double* host, uint64_t n) {
  return guarded([&] {
    check_vec(v);
    require(n == v->n && (host || n == 0), "size mismatch");
    if (n == 0) return;
    if (v->dtype == ALPS_ED_F64) {
      hip_check(hipMemcpyAsync(v->buf.p, host, n * 8, hipMemcpyHostToDevice, g_stream), "copy");
    } else {
      DevBuf tmp(n * 8);
      hip_check(hipMemcpyAsync(tmp.p, host, n * 8, hipMemcpyHostToDevice, g_stream), "copy");
      hip_check(launch_convert<float, double>(n, v->buf.as<float>(), tmp.as<double>(), nullptr, g_stream), "convert");
      hip_check(hipStreamSynchronize(g_stream), "sync");
    }
    hip_check(hipStreamSynchronize(g_stream), "sync");
  });
}

int alps_ed_vector_download(alps_ed_vector v, double* host, uint64_t n) {
  return guarded([&] {
    check_vec(v);
    require(n == v->n && (host || n == 0), "size mismatch");
    if (n == 0) return;
    if (v->dtype == ALPS_ED_F64) {
      hip_check(hipMemcpyAsync(host, v->buf.p, n * 8, hipMemcpyDeviceToHost, g_stream), "copy");
    } else {
      DevBuf tmp(n * 8);
      hip_check(launch_convert<double, float>(n, tmp.as<double>(), v->buf.as<float>(), nullptr, g_stream), "convert");
      hip_check(hipMemcpyAsync(host, tmp.p, n * 8, hipMemcpyDeviceToHost, g_stream), "copy");
      hip_check(hipStreamSynchronize(g_stream), "sync");
    }
    hip_check(hipStreamSynchronize(g_stream), "sync");
  });
}

int alps_ed_vector_fill_random(alps_ed_vector v, uint64_t seed, double lo, double hi) {
  return guarded([&] {
    check_vec(v);
    with_dtype(v->dtype, [&](auto z) {
      using V = decltype(z);
      hip_check(launch_fill_random<V>(v->n, v->buf.as<V>(), seed, lo, hi, g_stream), "fill_random");
    });
  });
}

int alps_ed_vector_fill(alps_ed_vector v, double value) {
  return guarded([&] {
    check_vec(v);
    with_dtype(v->dtype, [&](auto z) {
      using V = decltype(z);
      hip_check(launch_fill<V>(v->n, v->buf.as<V>(), value, g_stream), "fill");
    });
  });
}

int alps_ed_vector_copy(alps_ed_vector dst, alps_ed_vector src) {
  return guarded([&] {
    check_vec(dst);
    check_vec(src);
    require(dst->n == src->n, "size mismatch");
    if (dst == src) return;
    if (dst->dtype == src->dtype) {
      hip_check(hipMemcpyAsync(dst->buf.p, src->buf.p, src->n * dtype_size(src->dtype), hipMemcpyDeviceToDevice, g_stream), "copy");
    } else if (dst->dtype == ALPS_ED_F32) {
      hip_check(launch_convert<float, double>(src->n, dst->buf.as<float>(), src->buf.as<double>(), nullptr, g_stream), "convert");
    } else {
      hip_check(launch_convert<double, float>(src->n, dst->buf.as<double>(), src->buf.as<float>(), nullptr, g_stream), "convert");
    }
  });
}

int alps_ed_vector_scale(alps_ed_vector v, double a) {
  return guarded([&] {
    check_vec(v);
    with_dtype(v->dtype, [&](auto z) {
      using V = decltype(z);
      hip_check(launch_axpby<V>(v->n, v->buf.as<V>(), 0.0, nullptr, a, nullptr, nullptr, false, g_stream), "scale");
    });
  });
}

int alps_ed_vector_axpby(alps_ed_vector y, double a, alps_ed_vector x, double b) {
  return guarded([&] {
    check_vec(y);
    check_vec(x);
    require(x->n == y->n && x->dtype == y->dtype, "size or dtype mismatch");
    with_dtype(y->dtype, [&](auto z) {
      using V = decltype(z);
      hip_check(launch_axpby<V>(y->n, y->buf.as<V>(), a, x->buf.as<V>(), b, nullptr, nullptr, false, g_stream), "axpby");
    });
  });
}

int alps_ed_vector_dot(alps_ed_vector x, alps_ed_vector y, double* result) {
  return guarded([&] {
    check_vec(x);
    check_vec(y);
    require(result != nullptr, "null argument");
    require(x->n == y->n && x->dtype == y->dtype, "size or dtype mismatch");
    with_dtype(x->dtype, [&](auto z) {
      using V = decltype(z);
      *result = dot_sync<V>(x->n, x->buf.as<V>(), y->buf.as<V>());
    });
  });
}

int alps_ed_vector_nrm2(alps_ed_vector x, double* result) {
  return guarded([&] {
    check_vec(x);
    require(result != nullptr, "null argument");
    with_dtype(x->dtype, [&](auto z) {
      using V = decltype(z);
      *result = std::sqrt(dot_sync<V>(x->n, x->buf.as<V>(), x->buf.as<V>()));
    });
  });
}

int alps_ed_matvec(alps_ed_operator op, alps_ed_vector x, alps_ed_vector y) {
  return guarded([&] {
    require(op != nullptr, "operator is null");
    check_vec(x);
    check_vec(y);
    require(x->n == op->layout.D && y->n == op->layout.D, "vector size != operator dimension");
    require(x->dtype == y->dtype, "dtype mismatch");
    require(x != y, "in-place product is not supported");
    GridReduce none{};
    with_dtype(x->dtype, [&](auto z) {
      using V = decltype(z);
      hip_check(launch_hv<V>(op->dev, x->buf.as<V>(), y->buf.as<V>(), nullptr, none, g_stream), "hv");
    });
  });
}

// One make_step on caller vectors (lanczos.h:392-406): w = H v_cur,
// alpha = <v_cur, w>, w -= alpha v_cur + beta_prev v_prev, beta = |w|,
// w /= beta.  v_prev and v_cur are left unchanged.
int alps_ed_lanczos_step_vectors(alps_ed_operator op, alps_ed_vector v_prev, alps_ed_vector v_cur,
                                 alps_ed_vector w, double beta_prev, double* alpha, double* beta) {
  return guarded([&] {
    require(op && alpha && beta, "null argument");
    check_vec(v_prev);
    check_vec(v_cur);
    check_vec(w);
    const uint64_t n = op->layout.D;
    require(v_prev->n == n && v_cur->n == n && w->n == n, "vector size != operator dimension");
    require(v_prev->dtype == v_cur->dtype && v_cur->dtype == w->dtype, "dtype mismatch");
    Reducer& r = thread_reducer();
    r.ensure(std::max<size_t>(hv_blocks(op->dev), stream_blocks(n)));
    double* res = r.result(0);
    hip_check(hipMemcpyAsync(res + 2, &beta_prev, sizeof(double), hipMemcpyHostToDevice, g_stream), "copy");
    with_dtype(w->dtype, [&](auto z) {
      using V = decltype(z);
      V* wv = w->buf.as<V>();
      hip_check(launch_hv<V>(op->dev, v_cur->buf.as<V>(), wv, nullptr, r.make(0, res, nullptr, 0), g_stream), "hv");
      hip_check(launch_update<V>(n, wv, v_cur->buf.as<V>(), v_prev->buf.as<V>(), wv, res, nullptr, res + 2,
                                 nullptr, r.make(1, res + 1, res + 3, 1), g_stream),
                "update");
      hip_check(launch_axpby<V>(n, wv, 0.0, wv, 0.0, res + 3, nullptr, false, g_stream), "scale");
    });
    double h[2];
    hip_check(hipMemcpyAsync(h, res, 2 * sizeof(double), hipMemcpyDeviceToHost, g_stream), "copy");
    hip_check(hipStreamSynchronize(g_stream), "sync");
    *alpha = h[0];
    *beta = h[1];
  });
}

// ---------------------------------------------------------------------------
// device-resident Lanczos
// ---------------------------------------------------------------------------
int alps_ed_lanczos_create(alps_ed_operator op, alps_ed_dtype dtype, int max_steps, alps_ed_lanczos* out) {
  return guarded([&] {
    require(op && out, "null argument");
    require(dtype == ALPS_ED_F32 || dtype == ALPS_ED_F64, "unknown dtype");
    require(max_steps >= 1, "max_steps must be >= 1");
    require(op->layout.D >= 1, "empty sector");
    std::unique_ptr<alps_ed_lanczos_s> lz(new alps_ed_lanczos_s);
    lz->op = op;
    lz->dtype = dtype;
    lz->max_steps = max_steps;
    const uint64_t n = op->layout.D;
    const size_t es = dtype_size(dtype);
    lz->u[0].alloc(n * es);
    lz->u[1].alloc(n * es);
    lz->w.alloc(n * es);
    lz->alpha.alloc((max_steps + 2) * sizeof(double));
    lz->beta.alloc((max_steps + 2) * sizeof(double));
    lz->scale.alloc((max_steps + 3) * sizeof(double));
    lz->red.ensure(std::max<size_t>(hv_blocks(op->dev), stream_blocks(n)));
    lz->stream = g_stream;
    *out = lz.release();
  });
}

int alps_ed_lanczos_destroy(alps_ed_lanczos lz) {
  delete lz;
  return ALPS_ED_OK;
}

int alps_ed_lanczos_start(alps_ed_lanczos lz, uint64_t seed, alps_ed_vector start) {
  return guarded([&] {
    require(lz != nullptr, "null argument");
    const uint64_t n = lz->op->layout.D;
    lz->stream = g_stream;
    with_dtype(lz->dtype, [&](auto z) {
      using V = decltype(z);
      V* u0 = lz->u[0].as<V>();
      if (start) {
        require(start->n == n, "start vector size != operator dimension");
        if (start->dtype == lz->dtype)
          hip_check(hipMemcpyAsync(u0, start->buf.p, n * sizeof(V), hipMemcpyDeviceToDevice, lz->stream), "copy");
        else if (start->dtype == ALPS_ED_F64)
          hip_check(launch_convert<V, double>(n, u0, start->buf.as<double>(), nullptr, lz->stream), "convert");
        else
          hip_check(launch_convert<V, float>(n, u0, start->buf.as<float>(), nullptr, lz->stream), "convert");
      } else {
        hip_check(launch_fill_random<V>(n, u0, seed, -1.0, 1.0, lz->stream), "fill_random");
      }
      double* scale = lz->scale.as<double>();
      // scale[0] = 1/|u0|; the norm goes to a scratch slot of beta's tail
      hip_check(launch_dot<V>(n, u0, u0, lz->red.make(1, lz->beta.as<double>() + lz->max_steps + 1, scale, 1),
                              lz->stream),
                "norm");
      lanczos_step_dev<V>(lz, 0);
    });
    lz->steps = 1;
  });
}

int alps_ed_lanczos_steps(alps_ed_lanczos lz, int nsteps) {
  return guarded([&] {
    require(lz != nullptr, "null argument");
    if (lz->steps == 0) throw StateError("lanczos_steps called before lanczos_start");
    require(nsteps >= 0 && lz->steps + nsteps <= lz->max_steps, "steps beyond max_steps");
    with_dtype(lz->dtype, [&](auto z) {
      using V = decltype(z);
      for (int s = 0; s < nsteps; ++s) lanczos_step_dev<V>(lz, lz->steps + s);
    });
    lz->steps += nsteps;
  });
}

int alps_ed_lanczos_coefficients(alps_ed_lanczos lz, double* alpha, double* beta, int max_n, int* n) {
  return guarded([&] {
    require(lz && n, "null argument");
    const int m = std::min(max_n, lz->steps);
    if (m > 0) {
      require(alpha && beta, "null argument");
      hip_check(hipMemcpyAsync(alpha, lz->alpha.p, m * sizeof(double), hipMemcpyDeviceToHost, lz->stream), "copy");
      hip_check(hipMemcpyAsync(beta, lz->beta.p, m * sizeof(double), hipMemcpyDeviceToHost, lz->stream), "copy");
    }
    hip_check(hipStreamSynchronize(lz->stream), "sync");
    *n = lz->steps;
  });
}

int alps_ed_lanczos_vectors(alps_ed_lanczos lz, alps_ed_vector v_prev, alps_ed_vector v_cur) {
  return guarded([&] {
    require(lz != nullptr, "null argument");
    if (lz->steps == 0) throw StateError("no Lanczos vectors before lanczos_start");
    const uint64_t n = lz->op->layout.D;
    const int j = lz->steps;  // v_j is in u[j&1] with scale[j]
    alps_ed_vector outs[2] = {v_prev, v_cur};
    for (int q = 0; q < 2; ++q) {
      alps_ed_vector o = outs[q];
      if (!o) continue;
      require(o->n == n, "vector size != operator dimension");
      const int idx = j - 1 + q;
      const double* sc = lz->scale.as<double>() + idx;
      with_dtype(lz->dtype, [&](auto z) {
        using V = decltype(z);
        const V* src = lz->u[idx & 1].as<V>();
        if (o->dtype == ALPS_ED_F32)
          hip_check(launch_convert<float, V>(n, o->buf.as<float>(), src, sc, lz->stream), "convert");
        else
          hip_check(launch_convert<double, V>(n, o->buf.as<double>(), src, sc, lz->stream), "convert");
      });
    }
    hip_check(hipStreamSynchronize(lz->stream), "sync");
  });
}

// ---------------------------------------------------------------------------
// T-matrix analysis and the sparsediag solve
// ---------------------------------------------------------------------------
int alps_ed_tmatrix_analyze(int n, const double* alpha, const double* beta, double* eigenvalues,
                            double* errors, int* multiplicities, int* n_out) {
  return guarded([&] {
    require(n >= 1 && alpha && beta && n_out, "null argument or n < 1");
    double amax = alpha[0], bmax = beta[0];
    for (int i = 1; i < n; ++i) {
      amax = std::max(amax, alpha[i]);
      bmax = std::max(bmax, beta[i]);
    }
    ietl::tmatrix_analysis an;
    an.run(alpha, beta, n, amax, bmax, true);
    const int m = (int)an.eigval_noghost.size();
    for (int i = 0; i < m; ++i) {
      if (eigenvalues) eigenvalues[i] = an.eigval_noghost[i];
      if (errors) errors[i] = an.err_noghost[i];
      if (multiplicities) multiplicities[i] = an.mult_noghost[i];
    }
    *n_out = m;
  });
}

}  // extern "C"

namespace {

// lanczos_iteration_nlowest::converged on the prefix T[0..m) (iteration.h:116-130).
bool nlowest_converged(const std::vector<double>& a, const std::vector<double>& b, int m, int nev,
                       std::vector<double>* ev_out) {
  double amax = a[0], bmax = b[0];
  for (int i = 1; i < m; ++i) {
    amax = std::max(amax, a[i]);
    bmax = std::max(bmax, b[i]);
  }
  ietl::tmatrix_analysis an;
  an.run(a.data(), b.data(), m, amax, bmax, true);
  if (ev_out) *ev_out = an.eigval_noghost;
  const double tol = 100. * std::numeric_limits<double>::epsilon();
  if ((int)an.eigval_noghost.size() < nev) return false;
  for (int i = 0; i < nev; ++i)
    if (an.err_noghost[i] > std::max(tol, tol * std::fabs(an.eigval_noghost[i]))) return false;
  return true;
}

std::vector<double> tmatrix_eigenvalues(const std::vector<double>& a, const std::vector<double>& b, int m) {
  std::vector<double> ev;
  nlowest_converged(a, b, m, 1 << 30, &ev);
  return ev;
}

}  // namespace

extern "C" int alps_ed_sparsediag(const alps_ed_model* model, const alps_ed_solve_params* params,
                                  alps_ed_solve_result* result) {
  return guarded([&] {
    require(params && result, "null argument");
    auto t0 = std::chrono::steady_clock::now();
    std::memset(result, 0, sizeof(*result));
    std::unique_ptr<alps_ed_operator_s> op(new alps_ed_operator_s);
    op->layout = build_layout(model_spec(model));
    const uint64_t D = op->layout.D;
    result->dim = D;
    if (D == 0) return;  // sparsediag.h:101-102
    operator_upload(op.get());
    if (D == 1) {  // sparsediag.h:120-124: the single matrix element
      double v = 0.0;
      uint32_t c = 0;
      int cnt = 0;
      std::vector<uint32_t> cols(op->layout.bond_p1.size() + 1);
      std::vector<double> vals(cols.size());
      if (alps_ed_operator_elements(op.get(), 0, 1, (int)cols.size(), cols.data(), vals.data(), &cnt) != ALPS_ED_OK)
        throw DeviceError(g_last_error);
      for (int q = 0; q < cnt; ++q)
        if (cols[q] == c) v = vals[q];
      result->n_eigenvalues = 1;
      result->eigenvalues[0] = v;
      result->iterations = 0;
      return;
    }
    const int max_iter = params->max_iter > 0 ? params->max_iter : (int)std::min<uint64_t>(10 * D, 1000);
    const int nev = params->nev > 0 ? params->nev : 1;
    require(nev <= 64, "nev must be <= 64");
    alps_ed_lanczos lzh = nullptr;
    if (alps_ed_lanczos_create(op.get(), params->dtype, max_iter, &lzh) != ALPS_ED_OK) throw DeviceError(g_last_error);
    std::unique_ptr<alps_ed_lanczos_s> lz(lzh);
    if (alps_ed_lanczos_start(lz.get(), params->seed, nullptr) != ALPS_ED_OK) throw DeviceError(g_last_error);
    std::vector<double> a(max_iter + 1), b(max_iter + 1);
    int n = 1, checked = 1, final_n = -1;
    int error_code = 0;
    while (final_n < 0) {
      int target = params->fixed_iterations ? max_iter : std::min(max_iter, n + std::max(1, std::min(n / 8, 64)));
      if (target > n) {
        if (alps_ed_lanczos_steps(lz.get(), target - n) != ALPS_ED_OK) throw DeviceError(g_last_error);
        n = target;
      }
      int got = 0;
      if (alps_ed_lanczos_coefficients(lz.get(), a.data(), b.data(), n, &got) != ALPS_ED_OK)
        throw DeviceError(g_last_error);
      if (!params->fixed_iterations && n >= 2 && nlowest_converged(a, b, n, nev, nullptr)) {
        // first converged prefix in (checked, n], as the per-step test would find
        for (int m = std::max(2, checked + 1); m <= n; ++m)
          if (m == n || nlowest_converged(a, b, m, nev, nullptr)) {
            final_n = m;
            break;
          }
      } else if (n >= max_iter) {
        final_n = n;
        error_code = 1;  // iteration.h:80
      }
      checked = n;
    }
    std::vector<double> ev = tmatrix_eigenvalues(a, b, final_n);
    const int k = std::min<int>(nev, (int)ev.size());
    result->n_eigenvalues = k;
    for (int i = 0; i < k; ++i) result->eigenvalues[i] = ev[i];
    result->iterations = final_n;
    result->error_code = error_code;
    result->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  });
}
