// ed_capi.hip -- implementation of the C ABI declared in include/alps_ed.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cctype>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <map>
#include <new>
#include <thread>
#include <stdexcept>
#include <sstream>
#include <string>
#include <vector>

#include <ietl/tmatrix.h>

#include "alps_ed.h"
#include "ed_host.hpp"
#include "ed_kernels.hpp"
#include "ed_bloch.hpp"
#include "ed_measure.hpp"
#include "ed_sparsediag.hpp"
#include "ed_xml.hpp"

using namespace alps_ed;
using namespace alps_ed::host;

namespace {

// Scratch for reductions: two partial-sum arrays (one per producer kind,
// so a consumer never reads an array that is being rewritten) and result
// slots.  Partial counts follow the producing kernel's grid.
struct Reducer {
  DevBuf pa, pb, pc, results;
  size_t cap = 0;
  void ensure(size_t nblocks) {
    if (nblocks <= cap) return;
    pa.alloc(nblocks * sizeof(double));
    pb.alloc(nblocks * sizeof(double));
    pc.alloc(nblocks * sizeof(double));
    cap = nblocks;
    if (!results.p) results.alloc(16 * sizeof(double));
  }
  double* a() { return pa.as<double>(); }
  double* b() { return pb.as<double>(); }
  double* c() { return pc.as<double>(); }
  double* result(int i) { return results.as<double>() + i; }
};

Reducer& thread_reducer() {
  // one per thread, kept until the process ends: a thread-exit destructor
  // would free device memory during process exit, while the runtime's own
  // teardown may already be under way
  thread_local Reducer* r = new Reducer;
  r->ensure(2048);
  return *r;
}

}  // namespace

// ---------------------------------------------------------------------------
// object definitions
// ---------------------------------------------------------------------------

struct alps_ed_operator_s {
  // references: the caller's (released by alps_ed_operator_destroy) and one
  // per Lanczos object created on it (released by alps_ed_lanczos_destroy);
  // the operator is freed with the last
  std::atomic<int> refs{1};
  SectorLayout layout;
  DevLayout dev{};
  DevBuf tables;  // every layout table, packed (operator_upload)
  DevBuf lists;   // per-high-word lists built on the device: hh_off and xb_nb,
  DevBuf hh_ent;  // and the listed high-high flips
  DevBuf seg_rec; // per processing index: the segment's record (SegRec)
  uint64_t bytes = 0;
  uint64_t n = 0;  // vector length: D, or 2 * dk for a momentum sector (realified)
  int schedule = 0;  // H.v schedule for FP32 vectors (kHvSchedule)
  std::vector<double> schedule_seconds;  // per schedule, when tuned
  // momentum sector (ed_bloch.hpp): rows are the kept representatives
  bool bloch = false;
  DevBloch bdev{};
  DevBuf btables;  // the Bloch basis tables, packed (bloch_upload)
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
  bool keep_basis = false;
  // Lanczos vector u_j (v_j = scale[j] * u_j): ping-pong buffers, or, when the
  // basis is kept, chunks of kChunk consecutive vectors allocated on demand.
  static constexpr int kChunk = 32;
  DevBuf u[2], w;
  unsigned norm_count[2] = {0, 0};  // partials in each norm buffer
  std::vector<std::unique_ptr<PooledBuf>> chunks;
  DevBuf alpha, beta, scale;  // FP64 on device
  // host-mapped copies of alpha (first max_steps + 2) and beta (next
  // max_steps + 2), written by the kernels that form them
  MappedBuf mirror;
  double* alpha_h(int j) const { return static_cast<double*>(mirror.d) + j; }
  double* beta_h(int j) const { return static_cast<double*>(mirror.d) + max_steps + 2 + j; }
  const double* host_alpha() const { return static_cast<const double*>(mirror.h); }
  const double* host_beta() const { return static_cast<const double*>(mirror.h) + max_steps + 2; }
  Reducer red;
  hipStream_t stream = nullptr;
  // Partial reorthogonalization (ReorthDev, ed_kernels.hpp) of a kept
  // basis: the omega rows, step flags, Gram-Schmidt coefficients and
  // partials, and the device table of the kept vectors' addresses (mirrored
  // on the host; the mirror is reserved once, so queued copies from it stay
  // valid).
  bool reorth = false;
  DevBuf r_omega, r_flag, r_gamma, r_dot, r_hnorm, r_slots;
  std::vector<void*> slot_host;
  ReorthDev rdev{};

  // Steps queued past what a caller reads (a solve's speculative chunk) may
  // still run: they write the mapped mirror and the basis chunks, which go
  // back to the host heap and the buffer pool, so wait for them first.
  ~alps_ed_lanczos_s() {
    if (steps > 0) (void)hipStreamSynchronize(stream);
  }

  // slots are padded to 64 elements, so every slot is 256-byte aligned
  size_t vec_bytes() const { return ((op->n + 63) & ~uint64_t(63)) * (dtype == ALPS_ED_F32 ? 4 : 8); }
  void* slot(int j) {
    if (!keep_basis) return u[j & 1].p;
    const size_t c = (size_t)j / kChunk;
    while (chunks.size() <= c) {
      chunks.emplace_back(new PooledBuf(vec_bytes() * kChunk));
      if (reorth) publish_chunk(chunks.size() - 1);
    }
    return static_cast<char*>(chunks[c]->p) + (size_t)(j % kChunk) * vec_bytes();
  }
  // the addresses of chunk c's vectors into the device slot table
  void publish_chunk(size_t c) {
    const size_t first = c * kChunk;
    if (first >= slot_host.size()) return;
    const size_t cnt = std::min<size_t>(kChunk, slot_host.size() - first);
    for (size_t i = 0; i < cnt; ++i) slot_host[first + i] = static_cast<char*>(chunks[c]->p) + i * vec_bytes();
    hip_check(hipMemcpyAsync(static_cast<void**>(r_slots.p) + first, slot_host.data() + first, cnt * sizeof(void*),
                             hipMemcpyHostToDevice, stream),
              "slot table");
  }
};

namespace {

// H.v of an operator of either kind, with the number of partials its DOT
// form writes.
unsigned op_hv_blocks(const alps_ed_operator_s* op) { return op->bloch ? bloch_blocks(op->bdev) : hv_blocks(op->dev); }

template <class V>
hipError_t op_hv(const alps_ed_operator_s* op, const V* x, V* y, double* dot_out, hipStream_t st) {
  return op->bloch ? launch_bloch_hv<V>(op->dev, op->bdev, x, y, dot_out, st)
                   : launch_hv<V>(op->dev, x, y, dot_out, st, op->schedule);
}

// The H.v autotuner: an operator of a fixed-Sz sector with at least
// kTuneMinDim states times every FP32 schedule on its own sector (two
// round-robin passes of one warm-up and three timed launches each, HIP
// events on the engine's stream) and keeps the fastest.  Choices are cached per sector shape (sites, up
// spins, low-word size and the bonds), so later operators of the same shape
// reuse them.  Schedules differ only in speed, never in results.
constexpr uint64_t kTuneMinDim = 200000;
static_assert(ALPS_ED_HV_SCHEDULES == kHvSchedules, "the C header lists every schedule");

std::vector<uint64_t> schedule_key(const SectorLayout& L) {
  uint64_t h = 1469598103934665603ull;  // FNV-1a over the bond positions
  for (size_t b = 0; b < L.bond_p1.size(); ++b)
    for (uint32_t v : {L.bond_p1[b], L.bond_p2[b]}) h = (h ^ v) * 1099511628211ull;
  return {(uint64_t)L.N, (uint64_t)L.k, (uint64_t)L.B, L.ll.size(), L.hh.size(), L.xb.size(), h};
}

void tune_schedule(alps_ed_operator op) {
  if (op->bloch || op->layout.D < kTuneMinDim) return;
  static std::mutex mu;
  static std::map<std::vector<uint64_t>, std::pair<int, std::vector<double>>> cache;
  const std::vector<uint64_t> key = schedule_key(op->layout);
  {
    std::lock_guard<std::mutex> g(mu);
    auto it = cache.find(key);
    if (it != cache.end()) {
      op->schedule = it->second.first;
      op->schedule_seconds = it->second.second;
      return;
    }
  }
  const uint64_t n = op->layout.D;
  hipStream_t st = current_stream();
  DevBuf x(n * sizeof(float)), y(n * sizeof(float));
  hip_check(launch_fill_random<float>(n, x.as<float>(), 1, -1.0, 1.0, st), "fill_random");
  hipEvent_t e0, e1;
  hip_check(hipEventCreate(&e0), "event");
  hip_check(hipEventCreate(&e1), "event");
  // two round-robin passes, the faster of each schedule's two times: the
  // first launches of a pass see cold caches whatever the schedule
  std::vector<double> sec(kHvSchedules, 1e30);
  for (int pass = 0; pass < 2; ++pass)
    for (int s = 0; s < kHvSchedules; ++s) {
      hip_check(launch_hv<float>(op->dev, x.as<float>(), y.as<float>(), nullptr, st, s), "hv");
      hip_check(hipEventRecord(e0, st), "event");
      for (int r = 0; r < 3; ++r)
        hip_check(launch_hv<float>(op->dev, x.as<float>(), y.as<float>(), nullptr, st, s), "hv");
      hip_check(hipEventRecord(e1, st), "event");
      hip_check(hipEventSynchronize(e1), "event");
      float ms = 0.f;
      hip_check(hipEventElapsedTime(&ms, e0, e1), "event");
      sec[s] = std::min(sec[s], ms * 1e-3 / 3);
    }
  int best = 0;
  for (int s = 1; s < kHvSchedules; ++s)
    if (sec[s] < sec[best]) best = s;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  op->schedule = best;
  op->schedule_seconds = sec;
  std::lock_guard<std::mutex> g(mu);
  cache[key] = {best, sec};
}

// Host arrays packed into one device allocation with one copy: each array
// starts at a 256-byte boundary; `add` records an array and returns where it
// will live on the device once `upload` has run.
class PackedUpload {
 public:
  template <class T>
  const T** add(const std::vector<T>& v, const T** dst) {
    items_.push_back({v.data(), v.size() * sizeof(T), off_, reinterpret_cast<const void**>(dst)});
    off_ += (std::max<size_t>(v.size() * sizeof(T), 1) + 255) & ~size_t(255);
    return dst;
  }
  // one allocation into `buf`, one host-to-device copy; fills the pointers
  void upload(DevBuf& buf) {
    buf.alloc(std::max<size_t>(off_, 256));
    std::vector<char> stage(off_);
    for (const Item& it : items_) {
      if (it.bytes) std::memcpy(stage.data() + it.off, it.src, it.bytes);
      *it.dst = static_cast<char*>(buf.p) + it.off;
    }
    if (off_) hip_check(hipMemcpy(buf.p, stage.data(), off_, hipMemcpyHostToDevice), "upload");
  }
  size_t bytes() const { return off_; }

 private:
  struct Item {
    const void* src;
    size_t bytes, off;
    const void** dst;
  };
  std::vector<Item> items_;
  size_t off_ = 0;
};

void operator_upload(alps_ed_operator op) {
  SectorLayout& L = op->layout;
  DevLayout& d = op->dev;
  d = DevLayout{};
  d.N = L.N;
  d.k = L.k;
  d.B = L.B;
  d.NH = L.NH;
  d.D = (uint32_t)L.D;
  d.n_seg = (uint32_t)L.seg_hi.size();
  d.max_seg = L.max_seg;
  d.n_ll_p = L.n_ll_p;
  d.n_hh_p = L.n_hh_p;
  d.n_xb_p = L.n_xb_p;
  d.n_ll = (int)L.ll.size();
  d.n_hh = (int)L.hh.size();
  d.n_xb = (int)L.xb.size();
  d.n_bonds = (int)L.bond_p1.size();
  PackedUpload pk;
#define ALPS_ED_UP(field) pk.add(L.field, &d.field);
  ALPS_ED_UP(lo_rank)
  ALPS_ED_UP(lo_state)
  ALPS_ED_UP(lo_class_off)
  ALPS_ED_UP(hi_base)
  ALPS_ED_UP(seg_hi)
  ALPS_ED_UP(seg_start)
  ALPS_ED_UP(proc_hi)
  ALPS_ED_UP(diag_lo_cls)
  ALPS_ED_UP(diag_hi)
  ALPS_ED_UP(nbp16)
  ALPS_ED_UP(nbp_off)
  ALPS_ED_UP(xn16)
  ALPS_ED_UP(xn_off)
  ALPS_ED_UP(ll_t)
  ALPS_ED_UP(xb_c)
  const bool dev_lists = L.hh_off.empty();  // host_lists = false: built below
  const uint32_t *hh_mask = nullptr, *xb_hbit = nullptr;
  const double* hh_t = nullptr;
  if (dev_lists) {
    pk.add(L.hh_mask, &hh_mask);
    pk.add(L.hh_t, &hh_t);
    pk.add(L.xb_hbit, &xb_hbit);
  } else {
    ALPS_ED_UP(hh_off)
    ALPS_ED_UP(hh_ent)
    ALPS_ED_UP(xb_nb)
  }
  ALPS_ED_UP(site_dn)
  ALPS_ED_UP(site_up)
  ALPS_ED_UP(bond_p1)
  ALPS_ED_UP(bond_p2)
  ALPS_ED_UP(bond_par)
  ALPS_ED_UP(bond_anti)
  ALPS_ED_UP(bond_t)
#undef ALPS_ED_UP
  pk.upload(op->tables);
  op->bytes = pk.bytes();
  op->n = L.D;
  if (dev_lists && L.D > 0) {
    // count, prefix sums and cross-bond neighbours, then (once the entry
    // count is known) the entries; on the null stream, finished before return
    const uint32_t nhi = 1u << L.NH;
    const size_t off_bytes = ((size_t)(nhi + 1) * 4 + 255) & ~size_t(255);
    const size_t xb_bytes = std::max<size_t>((size_t)nhi * L.n_xb_p * 4, 4);
    op->lists.alloc(off_bytes + xb_bytes);
    uint32_t* off = op->lists.as<uint32_t>();
    uint32_t* xnb = reinterpret_cast<uint32_t*>(op->lists.as<char>() + off_bytes);
    hip_check(launch_hh_count(nhi, d.hi_base, hh_mask, (int)L.hh_mask.size(), off, xb_hbit, L.n_xb_p, xnb, nullptr),
              "high-high lists");
    uint32_t total = 0;
    hip_check(hipMemcpy(&total, off + nhi, sizeof(uint32_t), hipMemcpyDeviceToHost), "high-high lists");
    if (total == 0xffffffffu) throw std::invalid_argument("high-high flip lists exceed 2^32-1 entries");
    op->hh_ent.alloc(std::max<size_t>(total, 1) * sizeof(HHEntry));
    hip_check(launch_hh_fill(nhi, d.hi_base, hh_mask, hh_t, (int)L.hh_mask.size(), off, op->hh_ent.as<HHEntry>(),
                             nullptr),
              "high-high lists");
    hip_check(hipStreamSynchronize(nullptr), "high-high lists");
    d.hh_off = off;
    d.hh_ent = op->hh_ent.as<HHEntry>();
    d.xb_nb = xnb;
    op->bytes += op->lists.bytes + op->hh_ent.bytes;
  }
  if (d.n_seg > 0) {
    op->seg_rec.alloc((size_t)d.n_seg * sizeof(SegRec));
    hip_check(launch_seg_rec(d, op->seg_rec.as<SegRec>(), nullptr), "segment records");
    hip_check(hipStreamSynchronize(nullptr), "segment records");
    d.seg_rec = op->seg_rec.as<SegRec>();
    op->bytes += op->seg_rec.bytes;
  }
}

// Turns an uploaded operator into its momentum sector.
void bloch_upload(alps_ed_operator op, const BlochSector& b) {
  DevBloch& d = op->bdev;
  d = DevBloch{};
  d.dk = (uint32_t)b.rep_state.size();
  PackedUpload pk;
  pk.add(b.rep_state, &d.rep_state);
  pk.add(b.norm, &d.norm);
  pk.add(b.full_pk, &d.full_pk);
  pk.add(b.norm_tab, &d.norm_tab);
  pk.add(b.ph_re, &d.ph_re);
  pk.add(b.ph_im, &d.ph_im);
  // the flips of every row, listed once (the H.v streams them)
  const BlochFlipRows fr = bloch_flip_rows(op->layout, b);
  if (fr.ok) {
    pk.add(fr.row_ptr, &d.row_ptr);
    pk.add(fr.ent, &d.ent);
    pk.add(fr.diag, &d.diag);
    pk.add(fr.inv_norm, &d.inv_norm);
    pk.add(fr.coef, &d.coef);
    d.sh_c = fr.sh_c;
    d.mask_a = fr.mask_a;
  }
  pk.upload(op->btables);
  op->bloch = true;
  op->n = 2 * (uint64_t)b.rep_state.size();
  op->bytes += pk.bytes();
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
  hip_check(launch_dot<V>(n, x, y, r.a(), current_stream()), "dot");
  hip_check(launch_finish(Partials{r.a(), stream_blocks(n)}, nullptr, r.result(0), nullptr, kSum, current_stream()), "finish");
  double h = 0.0;
  hip_check(hipMemcpyAsync(&h, r.result(0), sizeof(double), hipMemcpyDeviceToHost, current_stream()), "copy");
  hip_check(hipStreamSynchronize(current_stream()), "sync");
  return h;
}

// One Lanczos step of the device-resident recurrence (see alps_ed_lanczos_*),
// two launches: w = H u_j with the partials of <u_j, w>; then the fused
// update, which forms scale_j = 1/|u_j| and beta_{j-1} from the norm
// partials of the previous stage, alpha_j = scale_j^2 <u_j, w>, and stores
// r = scale_j w - alpha_j v_j - beta_{j-1} v_{j-1} as u_{j+1} with its norm
// partials.  The norm partials alternate between two buffers: stage j (the
// start for j = 0, the update of step j-1 otherwise) wrote norm_buf(j).
// beta_j and scale_{j+1} of the last step are formed by finish_norm.
double* norm_buf(alps_ed_lanczos lz, int j) { return (j & 1) ? lz->red.c() : lz->red.b(); }

template <class V>
void lanczos_step_dev(alps_ed_lanczos lz, int j) {
  V* cur = static_cast<V*>(lz->slot(j));
  V* prev = j > 0 ? static_cast<V*>(lz->slot(j - 1)) : cur;
  V* next = static_cast<V*>(lz->slot(j + 1));
  V* w = lz->w.as<V>();
  double* alpha = lz->alpha.as<double>();
  double* beta = lz->beta.as<double>();
  double* scale = lz->scale.as<double>();
  hip_check(op_hv<V>(lz->op, cur, w, lz->red.a(), lz->stream), "hv");
  hip_check(launch_update_fused<V>(lz->op->n, w, cur, prev, next, Partials{lz->red.a(), op_hv_blocks(lz->op)},
                                   Partials{norm_buf(lz, j), lz->norm_count[j & 1]}, alpha + j,
                                   j > 0 ? beta + j - 1 : nullptr, scale + j, j > 0 ? scale + j - 1 : nullptr,
                                   norm_buf(lz, j + 1), lz->stream, lz->alpha_h(j), j > 0 ? lz->beta_h(j - 1) : nullptr),
            "update");
  lz->norm_count[(j + 1) & 1] = fused_blocks(lz->op->n);
  // partial reorthogonalization: the omega step of step j and, when it calls
  // for it, the Gram-Schmidt pass on u_{j+1} (replacing its norm partials)
  if (lz->reorth)
    hip_check(launch_reorth<V>(lz->op->n, j, next, scale, lz->rdev, norm_buf(lz, j + 1), lz->stream), "reorth");
}

// beta_{n-1} and scale_n after n steps (the host reads beta).
void finish_norm(alps_ed_lanczos lz, int n) {
  hip_check(launch_finish(Partials{norm_buf(lz, n), lz->norm_count[n & 1]}, nullptr, lz->beta.as<double>() + n - 1,
                          lz->scale.as<double>() + n, kNorm, lz->stream, lz->beta_h(n - 1)),
            "finish");
}

}  // namespace

// ---------------------------------------------------------------------------
// library
// ---------------------------------------------------------------------------
extern "C" {

int alps_ed_version(void) { return 100; }

const char* alps_ed_last_error(void) { return last_error().c_str(); }

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
  current_stream() = static_cast<hipStream_t>(s);
  return ALPS_ED_OK;
}

int alps_ed_synchronize(void) {
  return guarded([&] { hip_check(hipStreamSynchronize(current_stream()), "hipStreamSynchronize"); });
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
    tune_schedule(op.get());
    *out = op.release();
  });
}


int alps_ed_operator_create_momentum(const alps_ed_model* model, int n_translations, const int* perm,
                                     const double* phase_re, const double* phase_im, alps_ed_operator* out) {
  return guarded([&] {
    require(out != nullptr && model != nullptr, "null argument");
    const Translations t = read_translations(model->n_sites, n_translations, perm, phase_re, phase_im);
    std::unique_ptr<alps_ed_operator_s> op(new alps_ed_operator_s);
    op->layout = build_layout(model_spec(model));
    require(op->layout.D >= 1, "empty sector");
    operator_upload(op.get());
    bloch_upload(op.get(), bloch_sector(bloch_orbits(op->layout, t), t, op->layout.N));
    *out = op.release();
  });
}

int alps_ed_operator_schedule(alps_ed_operator op, int* schedule, double* seconds) {
  return guarded([&] {
    require(op != nullptr && schedule != nullptr, "null argument");
    *schedule = op->schedule;
    if (seconds)
      for (int s = 0; s < kHvSchedules; ++s)
        seconds[s] = op->schedule_seconds.empty() ? -1.0 : op->schedule_seconds[s];
  });
}

int alps_ed_operator_set_schedule(alps_ed_operator op, int schedule) {
  return guarded([&] {
    require(op != nullptr, "null argument");
    require(schedule >= 0 && schedule < kHvSchedules, "unknown schedule");
    op->schedule = schedule;
  });
}

int alps_ed_operator_destroy(alps_ed_operator op) {
  if (op && op->refs.fetch_sub(1) == 1) delete op;
  return ALPS_ED_OK;
}

int alps_ed_operator_dimension(alps_ed_operator op, uint64_t* dim) {
  return guarded([&] {
    require(op && dim, "null argument");
    *dim = op->n;
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
    require(!op->bloch, "explicit states are not available for a momentum sector");
    require(first + count <= op->layout.D, "state range beyond the dimension");
    if (count == 0) return;
    DevBuf d(count * sizeof(uint32_t));
    hip_check(launch_unrank(op->dev, first, count, d.as<uint32_t>(), current_stream()), "unrank");
    hip_check(hipMemcpyAsync(states_out, d.p, count * sizeof(uint32_t), hipMemcpyDeviceToHost, current_stream()), "copy");
    hip_check(hipStreamSynchronize(current_stream()), "sync");
  });
}

int alps_ed_operator_index(alps_ed_operator op, const uint32_t* states, uint64_t count, uint64_t* index_out) {
  return guarded([&] {
    require(op && ((states && index_out) || count == 0), "null argument");
    if (count == 0) return;
    DevBuf ds(count * sizeof(uint32_t)), di(count * sizeof(uint64_t));
    hip_check(hipMemcpyAsync(ds.p, states, count * sizeof(uint32_t), hipMemcpyHostToDevice, current_stream()), "copy");
    hip_check(launch_rank(op->dev, ds.as<uint32_t>(), count, di.as<uint64_t>(), current_stream()), "rank");
    hip_check(hipMemcpyAsync(index_out, di.p, count * sizeof(uint64_t), hipMemcpyDeviceToHost, current_stream()), "copy");
    hip_check(hipStreamSynchronize(current_stream()), "sync");
  });
}

int alps_ed_operator_elements(alps_ed_operator op, uint64_t row0, uint64_t nrows, int max_per_row,
                              uint32_t* cols, double* vals, int* counts) {
  return guarded([&] {
    require(op && ((cols && vals && counts) || nrows == 0), "null argument");
    require(!op->bloch, "explicit elements are not available for a momentum sector");
    require(row0 + nrows <= op->layout.D, "row range beyond the dimension");
    require(max_per_row >= (int)op->layout.bond_p1.size() + 1, "max_per_row must be >= n_bonds + 1");
    if (nrows == 0) return;
    const uint64_t ne = nrows * (uint64_t)max_per_row;
    DevBuf dc(ne * sizeof(uint32_t)), dv(ne * sizeof(double)), dn(nrows * sizeof(int));
    hip_check(launch_elements(op->dev, row0, nrows, max_per_row, dc.as<uint32_t>(), dv.as<double>(),
                              dn.as<int>(), current_stream()),
              "elements");
    hip_check(hipMemcpyAsync(cols, dc.p, ne * sizeof(uint32_t), hipMemcpyDeviceToHost, current_stream()), "copy");
    hip_check(hipMemcpyAsync(vals, dv.p, ne * sizeof(double), hipMemcpyDeviceToHost, current_stream()), "copy");
    hip_check(hipMemcpyAsync(counts, dn.p, nrows * sizeof(int), hipMemcpyDeviceToHost, current_stream()), "copy");
    hip_check(hipStreamSynchronize(current_stream()), "sync");
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
      hip_check(launch_fill<V>(n, p->buf.as<V>(), 0.0, current_stream()), "fill");
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
      hip_check(hipMemcpyAsync(v->buf.p, host, n * 8, hipMemcpyHostToDevice, current_stream()), "copy");
    } else {
      DevBuf tmp(n * 8);
      hip_check(hipMemcpyAsync(tmp.p, host, n * 8, hipMemcpyHostToDevice, current_stream()), "copy");
      hip_check(launch_convert<float, double>(n, v->buf.as<float>(), tmp.as<double>(), nullptr, current_stream()), "convert");
      hip_check(hipStreamSynchronize(current_stream()), "sync");
    }
    hip_check(hipStreamSynchronize(current_stream()), "sync");
  });
}

int alps_ed_vector_download(alps_ed_vector v, double* host, uint64_t n) {
  return guarded([&] {
    check_vec(v);
    require(n == v->n && (host || n == 0), "size mismatch");
    if (n == 0) return;
    if (v->dtype == ALPS_ED_F64) {
      hip_check(hipMemcpyAsync(host, v->buf.p, n * 8, hipMemcpyDeviceToHost, current_stream()), "copy");
    } else {
      DevBuf tmp(n * 8);
      hip_check(launch_convert<double, float>(n, tmp.as<double>(), v->buf.as<float>(), nullptr, current_stream()), "convert");
      hip_check(hipMemcpyAsync(host, tmp.p, n * 8, hipMemcpyDeviceToHost, current_stream()), "copy");
      hip_check(hipStreamSynchronize(current_stream()), "sync");
    }
    hip_check(hipStreamSynchronize(current_stream()), "sync");
  });
}

int alps_ed_vector_fill_random(alps_ed_vector v, uint64_t seed, double lo, double hi) {
  return guarded([&] {
    check_vec(v);
    with_dtype(v->dtype, [&](auto z) {
      using V = decltype(z);
      hip_check(launch_fill_random<V>(v->n, v->buf.as<V>(), seed, lo, hi, current_stream()), "fill_random");
    });
  });
}

int alps_ed_vector_fill(alps_ed_vector v, double value) {
  return guarded([&] {
    check_vec(v);
    with_dtype(v->dtype, [&](auto z) {
      using V = decltype(z);
      hip_check(launch_fill<V>(v->n, v->buf.as<V>(), value, current_stream()), "fill");
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
      hip_check(hipMemcpyAsync(dst->buf.p, src->buf.p, src->n * dtype_size(src->dtype), hipMemcpyDeviceToDevice, current_stream()), "copy");
    } else if (dst->dtype == ALPS_ED_F32) {
      hip_check(launch_convert<float, double>(src->n, dst->buf.as<float>(), src->buf.as<double>(), nullptr, current_stream()), "convert");
    } else {
      hip_check(launch_convert<double, float>(src->n, dst->buf.as<double>(), src->buf.as<float>(), nullptr, current_stream()), "convert");
    }
  });
}

int alps_ed_vector_scale(alps_ed_vector v, double a) {
  return guarded([&] {
    check_vec(v);
    with_dtype(v->dtype, [&](auto z) {
      using V = decltype(z);
      hip_check(launch_axpby<V>(v->n, v->buf.as<V>(), 0.0, nullptr, a, current_stream()), "scale");
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
      hip_check(launch_axpby<V>(y->n, y->buf.as<V>(), a, x->buf.as<V>(), b, current_stream()), "axpby");
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

int alps_ed_vector_spin_moments(alps_ed_operator op, alps_ed_vector v, double* norm2, double* sz, double* szsz) {
  return guarded([&] {
    require(op != nullptr, "operator is null");
    check_vec(v);
    require(!op->bloch, "spin sums need a fixed-Sz sector without momentum");
    require(v->n == op->layout.D, "vector size != operator dimension");
    require(op->layout.D >= 1, "empty sector");
    const DevLayout& L = op->dev;
    const int N = L.N;
    DevBuf stats((size_t)hv_blocks(L) * moment_stat_count(L) * sizeof(double));
    DevBuf out((size_t)(1 + N + N * N) * sizeof(double));
    hipStream_t st = current_stream();
    with_dtype(v->dtype, [&](auto z) {
      using V = decltype(z);
      hip_check(launch_spin_moments<V>(L, v->buf.as<V>(), stats.as<double>(), out.as<double>(), st), "moments");
    });
    std::vector<double> h(1 + N + N * N);
    hip_check(hipMemcpyAsync(h.data(), out.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, st), "copy");
    hip_check(hipStreamSynchronize(st), "sync");
    if (norm2) *norm2 = h[0];
    if (sz) std::copy(h.begin() + 1, h.begin() + 1 + N, sz);
    if (szsz) std::copy(h.begin() + 1 + N, h.end(), szsz);
  });
}

int alps_ed_vector_bloch_expand(alps_ed_operator op, alps_ed_vector w, alps_ed_vector re, alps_ed_vector im) {
  return guarded([&] {
    require(op != nullptr, "operator is null");
    check_vec(w);
    check_vec(re);
    check_vec(im);
    require(op->bloch, "the operator is not a momentum sector");
    require(w->n == op->n, "vector size != operator dimension");
    require(re->n == op->layout.D && im->n == op->layout.D, "outputs need the fixed-Sz sector's dimension");
    require(re->dtype == ALPS_ED_F64 && im->dtype == ALPS_ED_F64, "outputs must be FP64 vectors");
    require(re != im, "re and im must be different vectors");
    with_dtype(w->dtype, [&](auto z) {
      using V = decltype(z);
      hip_check(launch_bloch_expand<V>(op->bdev, (uint32_t)op->layout.D, w->buf.as<V>(), re->buf.as<double>(),
                                       im->buf.as<double>(), current_stream()),
                "bloch expand");
    });
  });
}

int alps_ed_vector_flip_pairs(alps_ed_operator op, alps_ed_vector v, int n_pairs, const int* site_i,
                              const int* site_j, double* out) {
  return guarded([&] {
    require(op != nullptr, "operator is null");
    check_vec(v);
    require(!op->bloch, "spin sums need a fixed-Sz sector without momentum");
    require(v->n == op->layout.D, "vector size != operator dimension");
    require(n_pairs >= 0 && (n_pairs == 0 || (site_i && site_j && out)), "null argument");
    if (n_pairs == 0) return;
    require(n_pairs <= (1 << 19), "at most 2^19 pairs per call");
    require(op->layout.D >= 1, "empty sector");
    const DevLayout& L = op->dev;
    std::vector<uint32_t> masks(2 * (size_t)n_pairs);
    for (int p = 0; p < n_pairs; ++p) {
      require(site_i[p] >= 0 && site_i[p] < L.N && site_j[p] >= 0 && site_j[p] < L.N, "site out of range");
      require(site_i[p] != site_j[p], "a flip pair needs two different sites");
      masks[2 * p] = 1u << (L.N - 1 - site_j[p]);
      masks[2 * p + 1] = 1u << (L.N - 1 - site_i[p]);
    }
    hipStream_t st = current_stream();
    DevBuf dm(masks.size() * sizeof(uint32_t));
    DevBuf part((size_t)n_pairs * flip_pair_blocks(L) * sizeof(double));
    DevBuf dout((size_t)n_pairs * sizeof(double));
    hip_check(hipMemcpyAsync(dm.p, masks.data(), masks.size() * sizeof(uint32_t), hipMemcpyHostToDevice, st),
              "copy");
    with_dtype(v->dtype, [&](auto z) {
      using V = decltype(z);
      hip_check(launch_flip_pairs<V>(L, v->buf.as<V>(), dm.as<uint32_t>(), n_pairs, part.as<double>(),
                                     dout.as<double>(), st),
                "flip pairs");
    });
    hip_check(hipMemcpyAsync(out, dout.p, (size_t)n_pairs * sizeof(double), hipMemcpyDeviceToHost, st), "copy");
    hip_check(hipStreamSynchronize(st), "sync");
  });
}

int alps_ed_matvec(alps_ed_operator op, alps_ed_vector x, alps_ed_vector y) {
  return guarded([&] {
    require(op != nullptr, "operator is null");
    check_vec(x);
    check_vec(y);
    require(x->n == op->n && y->n == op->n, "vector size != operator dimension");
    require(x->dtype == y->dtype, "dtype mismatch");
    require(x != y, "in-place product is not supported");
    with_dtype(x->dtype, [&](auto z) {
      using V = decltype(z);
      hip_check(op_hv<V>(op, x->buf.as<V>(), y->buf.as<V>(), nullptr, current_stream()), "hv");
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
    const uint64_t n = op->n;
    require(v_prev->n == n && v_cur->n == n && w->n == n, "vector size != operator dimension");
    require(v_prev->dtype == v_cur->dtype && v_cur->dtype == w->dtype, "dtype mismatch");
    Reducer& r = thread_reducer();
    r.ensure(std::max<size_t>(op_hv_blocks(op), stream_blocks(n)));
    double* res = r.result(0);
    hip_check(hipMemcpyAsync(res + 2, &beta_prev, sizeof(double), hipMemcpyHostToDevice, current_stream()), "copy");
    with_dtype(w->dtype, [&](auto z) {
      using V = decltype(z);
      V* wv = w->buf.as<V>();
      hip_check(op_hv<V>(op, v_cur->buf.as<V>(), wv, r.a(), current_stream()), "hv");
      hip_check(launch_finish(Partials{r.a(), op_hv_blocks(op)}, nullptr, res, nullptr, kAlpha, current_stream()), "finish");
      hip_check(launch_update<V>(n, wv, v_cur->buf.as<V>(), v_prev->buf.as<V>(), wv, res, nullptr, res + 2, nullptr,
                                 r.b(), current_stream()),
                "update");
      hip_check(launch_normalize<V>(n, wv, Partials{r.b(), stream_blocks(n)}, res + 1, current_stream()), "normalize");
    });
    double h[2];
    hip_check(hipMemcpyAsync(h, res, 2 * sizeof(double), hipMemcpyDeviceToHost, current_stream()), "copy");
    hip_check(hipStreamSynchronize(current_stream()), "sync");
    *alpha = h[0];
    *beta = h[1];
  });
}

}  // extern "C"

// Scratch of the allocation-free step: the H.v dot partials (one per H.v
// block), the norm partials of the update (one per streaming block) and the
// scalar results, 2 MiB + 8 KiB + 128 B.  Each GPU holds its own copy for the
// life of the process.
constexpr size_t kStaticHvPartials = size_t(1) << 18;
constexpr size_t kStaticNormPartials = 1024;
constexpr size_t kStaticResults = 16;
static_assert(kStaticNormPartials >= 1024, "stream_blocks() is at most 1024");
static __device__ __align__(256) double scratch[kStaticHvPartials + kStaticNormPartials + kStaticResults];

namespace {

// The device address of `scratch` on the current device, looked up once per
// device (the array's host-side name is not a device address).
double* static_scratch() {
  constexpr int kMaxDevices = 64;
  static std::mutex mu;
  static double* addr[kMaxDevices] = {};
  int dev = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  require(dev >= 0 && dev < kMaxDevices, "device ordinal beyond the static scratch table");
  std::lock_guard<std::mutex> g(mu);
  if (!addr[dev]) {
    void* p = nullptr;
    hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    addr[dev] = static_cast<double*>(p);
  }
  return addr[dev];
}

}  // namespace

extern "C" {

// alps_ed_lanczos_step_vectors with the reductions in the library's static
// device scratch: the same kernels in the same order, no device or host
// allocation.
int alps_ed_lanczos_step_vectors_noalloc(alps_ed_operator op, alps_ed_vector v_prev, alps_ed_vector v_cur,
                                         alps_ed_vector w, double beta_prev, double* alpha, double* beta) {
  return guarded([&] {
    require(op && alpha && beta, "null argument");
    require(op_hv_blocks(op) <= kStaticHvPartials,
            "operator too large for the static scratch (more than 262144 H.v blocks)");
    check_vec(v_prev);
    check_vec(v_cur);
    check_vec(w);
    const uint64_t n = op->n;
    require(v_prev->n == n && v_cur->n == n && w->n == n, "vector size != operator dimension");
    require(v_prev->dtype == v_cur->dtype && v_cur->dtype == w->dtype, "dtype mismatch");
    double* hv_part = static_scratch();
    double* norm_part = hv_part + kStaticHvPartials;
    double* res = norm_part + kStaticNormPartials;
    hipStream_t st = current_stream();
    // beta_prev enters through a one-element fill kernel, not a host copy
    hip_check(launch_fill<double>(1, res + 2, beta_prev, st), "fill");
    with_dtype(w->dtype, [&](auto z) {
      using V = decltype(z);
      V* wv = w->buf.as<V>();
      hip_check(op_hv<V>(op, v_cur->buf.as<V>(), wv, hv_part, st), "hv");
      hip_check(launch_finish(Partials{hv_part, op_hv_blocks(op)}, nullptr, res, nullptr, kAlpha, st), "finish");
      hip_check(launch_update<V>(n, wv, v_cur->buf.as<V>(), v_prev->buf.as<V>(), wv, res, nullptr, res + 2, nullptr,
                                 norm_part, st),
                "update");
      hip_check(launch_normalize<V>(n, wv, Partials{norm_part, stream_blocks(n)}, res + 1, st), "normalize");
    });
    double h[2];
    hip_check(hipMemcpyAsync(h, res, 2 * sizeof(double), hipMemcpyDeviceToHost, st), "copy");
    hip_check(hipStreamSynchronize(st), "sync");
    *alpha = h[0];
    *beta = h[1];
  });
}


// ---------------------------------------------------------------------------
// device-resident Lanczos
// ---------------------------------------------------------------------------
int alps_ed_lanczos_create(alps_ed_operator op, alps_ed_dtype dtype, int max_steps, int keep_basis,
                           alps_ed_lanczos* out) {
  return guarded([&] {
    require(op && out, "null argument");
    require(dtype == ALPS_ED_F32 || dtype == ALPS_ED_F64, "unknown dtype");
    require(max_steps >= 1, "max_steps must be >= 1");
    require(op->n >= 1, "empty sector");
    std::unique_ptr<alps_ed_lanczos_s> lz(new alps_ed_lanczos_s);
    lz->op = op;
    lz->dtype = dtype;
    lz->max_steps = max_steps;
    lz->keep_basis = keep_basis != 0;
    const uint64_t n = op->n;
    if (!lz->keep_basis) {
      lz->u[0].alloc(lz->vec_bytes());
      lz->u[1].alloc(lz->vec_bytes());
    }
    lz->w.alloc(lz->vec_bytes());
    lz->alpha.alloc((max_steps + 2) * sizeof(double));
    lz->beta.alloc((max_steps + 2) * sizeof(double));
    lz->mirror.alloc(2 * (size_t)(max_steps + 2) * sizeof(double));
    lz->scale.alloc((max_steps + 3) * sizeof(double));
    lz->red.ensure(std::max<size_t>(op_hv_blocks(op), stream_blocks(n)));
    lz->stream = current_stream();
    op->refs.fetch_add(1);
    *out = lz.release();
  });
}

int alps_ed_lanczos_destroy(alps_ed_lanczos lz) {
  if (!lz) return ALPS_ED_OK;
  alps_ed_operator_s* op = lz->op;
  delete lz;
  if (op && op->refs.fetch_sub(1) == 1) delete op;
  return ALPS_ED_OK;
}

int alps_ed_lanczos_set_reorthogonalization(alps_ed_lanczos lz, int mode) {
  return guarded([&] {
    require(lz != nullptr, "null argument");
    require(mode == 0 || mode == 1, "reorthogonalization mode must be 0 (none) or 1 (partial)");
    if (lz->steps != 0) throw StateError("the reorthogonalization mode is set before lanczos_start");
    if (mode == 0) {
      lz->reorth = false;
      return;
    }
    if (!lz->keep_basis) throw StateError("partial reorthogonalization needs a Lanczos object created with keep_basis");
    // the Gram-Schmidt update stages one coefficient per kept vector in LDS
    require(lz->max_steps <= 4096, "partial reorthogonalization supports at most 4096 steps");
    const uint32_t stride = (uint32_t)lz->max_steps + 2;
    lz->r_omega.alloc(3 * (size_t)stride * sizeof(double));
    lz->r_flag.alloc(stride * sizeof(int));
    lz->r_gamma.alloc(stride * sizeof(double));
    lz->r_dot.alloc((size_t)kReorthDotBlocks * stride * sizeof(double));
    lz->r_hnorm.alloc(stride * sizeof(double));
    lz->r_slots.alloc(stride * sizeof(void*));
    hip_check(hipMemsetAsync(lz->r_flag.p, 0, stride * sizeof(int), current_stream()), "memset");
    lz->slot_host.assign(stride, nullptr);
    lz->rdev = ReorthDev{lz->r_omega.as<double>(), stride, lz->r_flag.as<int>(), lz->r_gamma.as<double>(),
                         lz->r_dot.as<double>(), lz->r_hnorm.as<double>(), lz->alpha.as<double>(),
                         lz->beta.as<double>(),
                         static_cast<const void* const*>(lz->r_slots.p),
                         lz->dtype == ALPS_ED_F32 ? 0.5 * std::numeric_limits<float>::epsilon()
                                                  : 0.5 * std::numeric_limits<double>::epsilon()};
    lz->reorth = true;
    lz->stream = current_stream();
    for (size_t c = 0; c < lz->chunks.size(); ++c) lz->publish_chunk(c);
  });
}

int alps_ed_lanczos_reorthogonalizations(alps_ed_lanczos lz, int* count) {
  return guarded([&] {
    require(lz && count, "null argument");
    *count = 0;
    if (!lz->reorth || lz->steps == 0) return;
    std::vector<int> f(lz->steps);
    hip_check(hipMemcpyAsync(f.data(), lz->r_flag.p, f.size() * sizeof(int), hipMemcpyDeviceToHost, lz->stream),
              "copy");
    hip_check(hipStreamSynchronize(lz->stream), "sync");
    for (int v : f) *count += v != 0;
  });
}

int alps_ed_lanczos_start(alps_ed_lanczos lz, uint64_t seed, alps_ed_vector start) {
  return guarded([&] {
    require(lz != nullptr, "null argument");
    const uint64_t n = lz->op->n;
    lz->stream = current_stream();
    with_dtype(lz->dtype, [&](auto z) {
      using V = decltype(z);
      V* u0 = static_cast<V*>(lz->slot(0));
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
      // scale[0] = 1/|u0|; |u0| goes to a spare slot of beta
      hip_check(launch_dot<V>(n, u0, u0, norm_buf(lz, 0), lz->stream), "norm");
      lz->norm_count[0] = stream_blocks(n);
      hip_check(launch_finish(Partials{norm_buf(lz, 0), lz->norm_count[0]}, nullptr,
                              lz->beta.as<double>() + lz->max_steps + 1, lz->scale.as<double>(), kNorm, lz->stream),
                "finish");
      lanczos_step_dev<V>(lz, 0);
      finish_norm(lz, 1);
    });
    lz->steps = 1;
  });
}

int alps_ed_lanczos_steps(alps_ed_lanczos lz, int nsteps) {
  return guarded([&] {
    require(lz != nullptr, "null argument");
    if (lz->steps == 0) throw StateError("lanczos_steps called before lanczos_start");
    require(nsteps >= 0 && lz->steps + nsteps <= lz->max_steps, "steps beyond max_steps");
    if (nsteps == 0) return;
    with_dtype(lz->dtype, [&](auto z) {
      using V = decltype(z);
      for (int s = 0; s < nsteps; ++s) lanczos_step_dev<V>(lz, lz->steps + s);
      finish_norm(lz, lz->steps + nsteps);
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
    const uint64_t n = lz->op->n;
    const int j = lz->steps;  // v_j = scale[j] u_j
    alps_ed_vector outs[2] = {v_prev, v_cur};
    for (int q = 0; q < 2; ++q) {
      alps_ed_vector o = outs[q];
      if (!o) continue;
      require(o->n == n, "vector size != operator dimension");
      const int idx = j - 1 + q;
      const double* sc = lz->scale.as<double>() + idx;
      with_dtype(lz->dtype, [&](auto z) {
        using V = decltype(z);
        const V* src = static_cast<const V*>(lz->slot(idx));
        if (o->dtype == ALPS_ED_F32)
          hip_check(launch_convert<float, V>(n, o->buf.as<float>(), src, sc, lz->stream), "convert");
        else
          hip_check(launch_convert<double, V>(n, o->buf.as<double>(), src, sc, lz->stream), "convert");
      });
    }
    hip_check(hipStreamSynchronize(lz->stream), "sync");
  });
}

}  // extern "C"

namespace {

// Snapshot of a device recurrence (lanczos::save / load,
// src/ietl/lanczos.h:86-98, which archive the T matrix, startvector, vec2
// and n): everything a later make_step reads.  Layout: the header, then
// alpha, beta and scale (whole device arrays), the norm partials of the
// next stage, the saved Lanczos vectors (u_{n-1} and u_n, or the whole kept
// basis u_0..u_n), and with partial reorthogonalization the omega rows,
// step flags, Gram-Schmidt coefficients and |H| estimates.
struct LanczosSnapshotHeader {
  char magic[8];
  uint64_t model_hash;  // hash_model of the operator
  uint64_t n;           // vector length
  int32_t dtype, keep_basis, reorth, max_steps, steps;
  uint32_t norm_count;  // partials of the next stage
  uint64_t vec_bytes;   // bytes of one saved vector
  uint32_t first_vector, n_vectors;
  uint64_t bytes[4];    // alpha, beta, scale, norm partials
  uint64_t reorth_bytes[4];  // omega, flag, gamma, hnorm
};
constexpr char kSnapshotMagic[8] = {'A', 'L', 'P', 'S', 'L', 'Z', 'S', '1'};

// FNV-1a over the model the operator was built from (sites, up spins, the
// reference-order site and bond terms) and the vector length.
uint64_t hash_model(const alps_ed_operator_s* op) {
  uint64_t h = 1469598103934665603ull;
  auto mix = [&](const void* p, size_t len) {
    const unsigned char* c = static_cast<const unsigned char*>(p);
    for (size_t i = 0; i < len; ++i) h = (h ^ c[i]) * 1099511628211ull;
  };
  const SectorLayout& L = op->layout;
  mix(&L.N, sizeof L.N);
  mix(&L.k, sizeof L.k);
  mix(&op->n, sizeof op->n);
  const int bl = op->bloch ? 1 : 0;
  mix(&bl, sizeof bl);
  auto vec = [&](const auto& v) { mix(v.data(), v.size() * sizeof(v[0])); };
  vec(L.site_dn);
  vec(L.site_up);
  vec(L.bond_p1);
  vec(L.bond_p2);
  vec(L.bond_par);
  vec(L.bond_anti);
  vec(L.bond_t);
  return h;
}

LanczosSnapshotHeader snapshot_header(const alps_ed_lanczos_s* lz) {
  LanczosSnapshotHeader hd{};
  std::memcpy(hd.magic, kSnapshotMagic, 8);
  hd.model_hash = hash_model(lz->op);
  hd.n = lz->op->n;
  hd.dtype = lz->dtype;
  hd.keep_basis = lz->keep_basis;
  hd.reorth = lz->reorth;
  hd.max_steps = lz->max_steps;
  hd.steps = lz->steps;
  hd.norm_count = lz->steps > 0 ? lz->norm_count[lz->steps & 1] : 0u;
  hd.vec_bytes = lz->op->n * (lz->dtype == ALPS_ED_F32 ? 4 : 8);
  hd.first_vector = lz->steps == 0 ? 0u : (lz->keep_basis ? 0u : (uint32_t)(lz->steps - 1));
  hd.n_vectors = lz->steps == 0 ? 0u : (uint32_t)(lz->steps + 1) - hd.first_vector;
  hd.bytes[0] = lz->alpha.bytes;
  hd.bytes[1] = lz->beta.bytes;
  hd.bytes[2] = lz->scale.bytes;
  hd.bytes[3] = (uint64_t)hd.norm_count * sizeof(double);
  if (lz->reorth) {
    hd.reorth_bytes[0] = lz->r_omega.bytes;
    hd.reorth_bytes[1] = lz->r_flag.bytes;
    hd.reorth_bytes[2] = lz->r_gamma.bytes;
    hd.reorth_bytes[3] = lz->r_hnorm.bytes;
  }
  return hd;
}

uint64_t snapshot_bytes(const LanczosSnapshotHeader& hd) {
  uint64_t b = sizeof hd + hd.bytes[0] + hd.bytes[1] + hd.bytes[2] + hd.bytes[3] + hd.n_vectors * hd.vec_bytes;
  for (uint64_t r : hd.reorth_bytes) b += r;
  return b;
}

}  // namespace

extern "C" {

int alps_ed_lanczos_snapshot_size(alps_ed_lanczos lz, size_t* bytes) {
  return guarded([&] {
    require(lz && bytes, "null argument");
    *bytes = (size_t)snapshot_bytes(snapshot_header(lz));
  });
}

int alps_ed_lanczos_save(alps_ed_lanczos lz, void* buf, size_t len) {
  return guarded([&] {
    require(lz && buf, "null argument");
    const LanczosSnapshotHeader hd = snapshot_header(lz);
    require(len >= snapshot_bytes(hd), "snapshot buffer smaller than alps_ed_lanczos_snapshot_size");
    char* o = static_cast<char*>(buf);
    std::memcpy(o, &hd, sizeof hd);
    o += sizeof hd;
    auto get = [&](const void* src, uint64_t b) {
      if (b) hip_check(hipMemcpyAsync(o, src, b, hipMemcpyDeviceToHost, lz->stream), "snapshot copy");
      o += b;
    };
    get(lz->alpha.p, hd.bytes[0]);
    get(lz->beta.p, hd.bytes[1]);
    get(lz->scale.p, hd.bytes[2]);
    if (hd.norm_count) get(norm_buf(lz, lz->steps), hd.bytes[3]);
    for (uint32_t v = 0; v < hd.n_vectors; ++v) get(lz->slot((int)(hd.first_vector + v)), hd.vec_bytes);
    if (hd.reorth) {
      get(lz->r_omega.p, hd.reorth_bytes[0]);
      get(lz->r_flag.p, hd.reorth_bytes[1]);
      get(lz->r_gamma.p, hd.reorth_bytes[2]);
      get(lz->r_hnorm.p, hd.reorth_bytes[3]);
    }
    hip_check(hipStreamSynchronize(lz->stream), "sync");
  });
}

int alps_ed_lanczos_load(alps_ed_lanczos lz, const void* buf, size_t len) {
  return guarded([&] {
    require(lz && buf, "null argument");
    require(len >= sizeof(LanczosSnapshotHeader), "snapshot too short");
    LanczosSnapshotHeader hd;
    std::memcpy(&hd, buf, sizeof hd);
    require(std::memcmp(hd.magic, kSnapshotMagic, 8) == 0, "not a Lanczos snapshot");
    require(len >= snapshot_bytes(hd), "snapshot truncated");
    if (lz->steps != 0) throw StateError("alps_ed_lanczos_load needs a Lanczos object before lanczos_start");
    require(hd.model_hash == hash_model(lz->op) && hd.n == lz->op->n, "snapshot of a different operator");
    require(hd.dtype == lz->dtype && hd.keep_basis == (int)lz->keep_basis && hd.max_steps == lz->max_steps,
            "snapshot of a Lanczos object with another dtype, keep_basis or max_steps");
    require(hd.reorth == (int)lz->reorth, "snapshot with another reorthogonalization mode");
    require(hd.steps >= 1 && hd.steps <= hd.max_steps, "snapshot has no steps");
    const LanczosSnapshotHeader mine = snapshot_header(lz);  // this object's array sizes
    require(hd.bytes[0] == mine.bytes[0] && hd.bytes[1] == mine.bytes[1] && hd.bytes[2] == mine.bytes[2] &&
                hd.vec_bytes == mine.vec_bytes && hd.reorth_bytes[0] == mine.reorth_bytes[0] &&
                hd.reorth_bytes[1] == mine.reorth_bytes[1] && hd.reorth_bytes[2] == mine.reorth_bytes[2] &&
                hd.reorth_bytes[3] == mine.reorth_bytes[3],
            "snapshot layout does not match this Lanczos object");
    require(hd.norm_count <= lz->red.cap, "snapshot partials exceed this object's reduction buffers");
    const uint32_t want_first = hd.keep_basis ? 0u : (uint32_t)(hd.steps - 1);
    require(hd.first_vector == want_first && hd.n_vectors == (uint32_t)(hd.steps + 1) - want_first,
            "snapshot vector set does not match its step count");
    lz->stream = current_stream();
    const char* in = static_cast<const char*>(buf) + sizeof hd;
    auto put = [&](void* dst, uint64_t b) {
      if (b) hip_check(hipMemcpyAsync(dst, in, b, hipMemcpyHostToDevice, lz->stream), "snapshot copy");
      in += b;
    };
    const char* alpha_in = in;
    put(lz->alpha.p, hd.bytes[0]);
    const char* beta_in = in;
    put(lz->beta.p, hd.bytes[1]);
    put(lz->scale.p, hd.bytes[2]);
    if (hd.norm_count) put(norm_buf(lz, hd.steps), hd.bytes[3]);
    for (uint32_t v = 0; v < hd.n_vectors; ++v) put(lz->slot((int)(hd.first_vector + v)), hd.vec_bytes);
    if (hd.reorth) {
      put(lz->r_omega.p, hd.reorth_bytes[0]);
      put(lz->r_flag.p, hd.reorth_bytes[1]);
      put(lz->r_gamma.p, hd.reorth_bytes[2]);
      put(lz->r_hnorm.p, hd.reorth_bytes[3]);
    }
    hip_check(hipStreamSynchronize(lz->stream), "sync");
    // the host mirror of alpha and beta, as the kernels that formed them
    // would have written it
    double* mh = static_cast<double*>(lz->mirror.h);
    const size_t cnt = (size_t)lz->max_steps + 2;
    std::memcpy(mh, alpha_in, std::min<uint64_t>(hd.bytes[0], cnt * sizeof(double)));
    std::memcpy(mh + cnt, beta_in, std::min<uint64_t>(hd.bytes[1], cnt * sizeof(double)));
    lz->norm_count[hd.steps & 1] = hd.norm_count;
    lz->steps = hd.steps;
  });
}

int alps_ed_lanczos_ritz(alps_ed_lanczos lz, int nvec, int m, const double* coeffs, alps_ed_vector* outs) {
  return guarded([&] {
    require(lz && coeffs && outs, "null argument");
    if (!lz->keep_basis) throw StateError("Ritz vectors need a Lanczos object created with keep_basis");
    require(m >= 1 && m <= lz->steps, "m must be in [1, steps]");
    require(nvec >= 0, "nvec must be >= 0");
    const uint64_t n = lz->op->n;
    for (int i = 0; i < nvec; ++i) {
      check_vec(outs[i]);
      require(outs[i]->n == n && outs[i]->dtype == ALPS_ED_F64, "Ritz outputs must be FP64 of the dimension");
    }
    std::vector<void*> slots(m);
    for (int j = 0; j < m; ++j) slots[j] = lz->slot(j);
    DevBuf dslots(m * sizeof(void*)), dcoef((size_t)kMaxRitz * m * sizeof(double)), douts(kMaxRitz * sizeof(void*));
    hip_check(hipMemcpyAsync(dslots.p, slots.data(), m * sizeof(void*), hipMemcpyHostToDevice, lz->stream), "copy");
    for (int i0 = 0; i0 < nvec; i0 += kMaxRitz) {
      const int nv = std::min(kMaxRitz, nvec - i0);
      std::vector<void*> op(kMaxRitz, nullptr);
      for (int i = 0; i < nv; ++i) op[i] = outs[i0 + i]->buf.p;
      hip_check(hipMemcpyAsync(dcoef.p, coeffs + (size_t)i0 * m, (size_t)nv * m * sizeof(double),
                               hipMemcpyHostToDevice, lz->stream), "copy");
      hip_check(hipMemcpyAsync(douts.p, op.data(), kMaxRitz * sizeof(void*), hipMemcpyHostToDevice, lz->stream), "copy");
      with_dtype(lz->dtype, [&](auto z) {
        using V = decltype(z);
        hip_check(launch_ritz<V>(n, dslots.as<const V*>(), lz->scale.as<double>(), m, nv, dcoef.as<double>(),
                                 douts.as<double*>(), lz->stream),
                  "ritz");
      });
      hip_check(hipStreamSynchronize(lz->stream), "sync");  // buffers are reused by the next pass
    }
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

void check_status(int st) {
  if (st != ALPS_ED_OK) throw DeviceError(last_error());
}

// Steps j < upto whose new vector u_{j+1} was reorthogonalized.
int reorthogonalized_steps(alps_ed_lanczos lz, int upto) {
  upto = std::min(upto, lz->steps);
  if (!lz->reorth || upto <= 0) return 0;
  std::vector<int> f(upto);
  hip_check(hipMemcpyAsync(f.data(), lz->r_flag.p, f.size() * sizeof(int), hipMemcpyDeviceToHost, lz->stream), "copy");
  hip_check(hipStreamSynchronize(lz->stream), "sync");
  int c = 0;
  for (int v : f) c += v != 0;
  return c;
}

// Ritz vectors (FP64, device) of the T[0..m) eigenvalues at the given
// positions, from the stored basis.
// `evals` are the eigenvalues of T[0..m) in ascending order, as
// analyze_prefix returns them; each eigenvector of T comes from inverse
// iteration at its eigenvalue (O(m) per vector instead of the full O(m^3)
// eigendecomposition).
std::vector<std::unique_ptr<alps_ed_vector_s>> ritz_vectors(alps_ed_lanczos lz, const std::vector<double>& a,
                                                            const std::vector<double>& b, int m,
                                                            const std::vector<int>& index,
                                                            const std::vector<double>& evals) {
  const uint64_t n = lz->op->n;
  std::vector<std::unique_ptr<alps_ed_vector_s>> out;
  for (size_t i0 = 0; i0 < index.size(); i0 += kMaxRitz) {
    const int nv = (int)std::min<size_t>(kMaxRitz, index.size() - i0);
    std::vector<double> coef((size_t)nv * m);
    for (int i = 0; i < nv; ++i) {
      require(index[i0 + i] >= 0 && index[i0 + i] < (int)evals.size(), "Ritz index beyond the T spectrum");
      tridiag_eigenvector(a.data(), b.data(), m, evals[index[i0 + i]], coef.data() + (size_t)i * m);
    }
    std::vector<alps_ed_vector> ys(nv, nullptr);
    for (int i = 0; i < nv; ++i) {
      check_status(alps_ed_vector_create(n, ALPS_ED_F64, &ys[i]));
      out.emplace_back(ys[i]);
    }
    check_status(alps_ed_lanczos_ritz(lz, nv, m, coef.data(), ys.data()));
  }
  return out;
}

// Rayleigh quotient y.Hy / y.y (FP64 accumulation) of a device FP64 vector.
double rayleigh_quotient(alps_ed_lanczos lz, alps_ed_vector y, alps_ed_vector hy) {
  const uint64_t n = lz->op->n;
  double* res = lz->red.result(0);
  hip_check(op_hv<double>(lz->op, y->buf.as<double>(), hy->buf.as<double>(), lz->red.a(), lz->stream), "hv");
  hip_check(launch_finish(Partials{lz->red.a(), op_hv_blocks(lz->op)}, nullptr, res, nullptr, kSum, lz->stream),
            "finish");
  hip_check(launch_dot<double>(n, y->buf.as<double>(), y->buf.as<double>(), lz->red.b(), lz->stream), "dot");
  hip_check(launch_finish(Partials{lz->red.b(), stream_blocks(n)}, nullptr, res + 1, nullptr, kSum, lz->stream),
            "finish");
  double h[2];
  hip_check(hipMemcpyAsync(h, res, 2 * sizeof(double), hipMemcpyDeviceToHost, lz->stream), "copy");
  hip_check(hipStreamSynchronize(lz->stream), "sync");
  return h[0] / h[1];
}

// Rayleigh quotients of the Ritz vectors of T[0..m) eigenvalues with the
// given positions, from the stored FP32 basis; the vectors are handed to
// `keep` (in the same order) when it is set.
std::vector<double> refine_ritz(alps_ed_lanczos lz, const std::vector<double>& a, const std::vector<double>& b,
                                int m, const std::vector<int>& index, const std::vector<double>& evals,
                                std::vector<std::unique_ptr<alps_ed_vector_s>>* keep) {
  std::vector<std::unique_ptr<alps_ed_vector_s>> ys = ritz_vectors(lz, a, b, m, index, evals);
  alps_ed_vector hy = nullptr;
  check_status(alps_ed_vector_create(lz->op->n, ALPS_ED_F64, &hy));
  std::unique_ptr<alps_ed_vector_s> hold_hy(hy);
  std::vector<double> rq;
  for (auto& y : ys) rq.push_back(rayleigh_quotient(lz, y.get(), hy));
  if (keep)
    for (auto& y : ys) keep->push_back(std::move(y));
  return rq;
}

// One sector as SparseDiagMatrix::do_subspace solves it (sparsediag.h:93-149),
// keeping the operator and, when `vectors` is set, the eigenvectors of the
// reported eigenvalues (FP64 Ritz vectors from the kept basis, as
// lanczos::eigenvectors provides them for the measurements, :125-144).
struct SectorSolve {
  std::unique_ptr<alps_ed_operator_s> op;
  alps_ed_solve_result r{};
  std::vector<std::unique_ptr<alps_ed_vector_s>> vectors;
};

// A momentum sector of a fixed-Sz sector: its layout, orbits and the
// translations at the momentum (the first must be the identity).
struct MomentumInput {
  const SectorLayout* layout;
  const BlochOrbits* orbits;
  const Translations* trans;
};

SectorSolve solve_sector(const alps_ed_model* model, const alps_ed_solve_params* params, bool want_vectors,
                         const MomentumInput* mom = nullptr) {
  auto t0 = std::chrono::steady_clock::now();
  SectorSolve out;
  alps_ed_solve_result* result = &out.r;
  std::memset(result, 0, sizeof(*result));
  out.op.reset(new alps_ed_operator_s);
  alps_ed_operator_s* op = out.op.get();
  op->layout = mom ? *mom->layout : build_layout(model_spec(model));
  if (op->layout.D > 0) operator_upload(op);
  if (mom && op->layout.D > 0) bloch_upload(op, bloch_sector(*mom->orbits, *mom->trans, op->layout.N));
  const uint64_t D = mom ? op->n / 2 : op->layout.D;  // the sector's dimension
  result->dim = D;
  auto seconds = [&] { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); };
  if (D == 0) return out;  // sparsediag.h:101-102
  if (D == 1 && mom) {  // the single (real) element of a one-state momentum sector
    alps_ed_vector x = nullptr, y = nullptr;
    check_status(alps_ed_vector_create(2, ALPS_ED_F64, &x));
    std::unique_ptr<alps_ed_vector_s> hx(x);
    check_status(alps_ed_vector_create(2, ALPS_ED_F64, &y));
    std::unique_ptr<alps_ed_vector_s> hy(y);
    const double e0[2] = {1.0, 0.0};
    check_status(alps_ed_vector_upload(x, e0, 2));
    check_status(alps_ed_matvec(op, x, y));
    double h[2];
    check_status(alps_ed_vector_download(y, h, 2));
    result->n_eigenvalues = 1;
    result->eigenvalues[0] = h[0];
    if (want_vectors) out.vectors.push_back(std::move(hx));
    result->seconds = seconds();
    return out;
  }
  if (D == 1) {  // sparsediag.h:120-124: the single matrix element
    std::vector<uint32_t> cols(op->layout.bond_p1.size() + 1);
    std::vector<double> vals(cols.size());
    int cnt = 0;
    check_status(alps_ed_operator_elements(op, 0, 1, (int)cols.size(), cols.data(), vals.data(), &cnt));
    double v = 0.0;
    for (int q = 0; q < cnt; ++q)
      if (cols[q] == 0) v = vals[q];
    result->n_eigenvalues = 1;
    result->eigenvalues[0] = v;
    if (want_vectors) {
      alps_ed_vector one = nullptr;
      check_status(alps_ed_vector_create(1, ALPS_ED_F64, &one));
      out.vectors.emplace_back(one);
      check_status(alps_ed_vector_fill(one, 1.0));
    }
    result->seconds = seconds();
    return out;
  }
  // FP32 storage for any number of levels; partial reorthogonalization where
  // copies of converged levels would delay higher ones (kFp32MaxNev)
  const bool fp32 = use_fp32_reorth(params->dtype == ALPS_ED_F32, D);
  if (fp32) tune_schedule(op);
  const int max_iter = params->max_iter > 0 ? params->max_iter : (int)std::min<uint64_t>(10 * D, 1000);
  const int nev = params->nev > 0 ? params->nev : 1;
  require(nev <= 64, "nev must be <= 64");
  alps_ed_lanczos lzh = nullptr;
  check_status(
      alps_ed_lanczos_create(op, fp32 ? ALPS_ED_F32 : ALPS_ED_F64, max_iter, (fp32 || want_vectors) ? 1 : 0, &lzh));
  std::unique_ptr<alps_ed_lanczos_s> lz(lzh);
  if (fp32 && nev > kFp32MaxNev) check_status(alps_ed_lanczos_set_reorthogonalization(lz.get(), 1));
  check_status(alps_ed_lanczos_start(lz.get(), params->seed, nullptr));
  std::vector<double> a(max_iter + 1), b(max_iter + 1);
  int final_n = -1, error_code = 0;
  {
    // alpha and beta reach the host through the Lanczos object's mapped
    // mirror, which the kernels write as they form them: after each chunk of
    // steps an event is queued, and the next chunk is queued before the host
    // waits for the previous one, so the stopping test of prefix n runs while
    // the device computes the next steps.  Steps past the stopping point are
    // discarded (the reported T matrix and the Ritz vectors use prefix
    // final_n only).
    // Prefixes queued and not yet tested, each with the event that marks
    // its completion (at most two: the one being waited for and the next).
    Event evs[3];
    struct Pending {
      int n;
      hipEvent_t e;
    };
    std::vector<Pending> pending;
    int queued = 1, next_ev = 0;
    auto queue_to = [&](int to) {
      if (to > queued) check_status(alps_ed_lanczos_steps(lz.get(), to - queued));
      queued = std::max(queued, to);
      hipEvent_t e = evs[next_ev].e;
      next_ev = (next_ev + 1) % 3;
      hip_check(hipEventRecord(e, lz->stream), "event");
      pending.push_back({to, e});
    };
    // The rule is tested at least eight steps apart, at most 64 (an eighth
    // of the length in between); when it holds, the first prefix at which
    // it holds is found by scanning forward, so the T matrix is the one the
    // per-step test of the reference stops at.  may_be_converged skips
    // prefixes where the rule cannot hold without the full analysis.  For
    // one level, the stopping step is predicted from the decay of the lowest
    // Ritz value's error bound between the last two tests, and the chunk
    // queued ahead ends just past it (fewer steps computed and discarded).
    int predicted = -1;  // predicted stopping step, -1 when unknown
    auto next_target = [&](int from) {
      if (params->fixed_iterations) return max_iter;
      int to = from + std::max(8, std::min(from / 8, 64));
      if (predicted > 0) to = std::min(to, std::max(from + 2, predicted + 1));
      return std::min(max_iter, to);
    };
    auto holds = [&](int m) {
      return may_be_converged(a.data(), b.data(), m, fp32) &&
             nlowest_converged(stop_analysis(a, b, m, fp32, nev), nev, fp32);
    };
    int checked = 1, err_n = -1;
    double err_prev = 0.0;
    queue_to(1);
    while (final_n < 0) {
      const int n = pending.front().n;
      // speculation: the next chunk is queued before the host waits for
      // prefix n, unless the prediction puts the stop at or before n
      if (pending.size() == 1 && queued < max_iter && !(predicted > 0 && predicted <= n)) queue_to(next_target(queued));
      hip_check(hipEventSynchronize(pending.front().e), "event");
      pending.erase(pending.begin());
      std::copy(lz->host_alpha(), lz->host_alpha() + n, a.begin());
      std::copy(lz->host_beta(), lz->host_beta() + n, b.begin());
      // A zero beta_j ends the recurrence: T[0..j+1) is then exact (an
      // invariant subspace); later coefficients would divide by it.  Other
      // non-finite coefficients (NaN or Inf in the start vector or the
      // couplings) are an error, reported instead of analysed.
      int good = n;
      bool exact = false;
      for (int j = 0; j < n; ++j) {
        if (!std::isfinite(a[j]) || !std::isfinite(b[j]))
          throw std::runtime_error("non-finite Lanczos coefficient at step " + std::to_string(j));
        if (b[j] == 0.0) {
          good = j + 1;
          exact = true;
          break;
        }
      }
      if (!params->fixed_iterations && good >= 2 && holds(good))
        for (int m = std::max(2, checked + 1); m <= good; ++m)
          if (m == good || holds(m)) {
            final_n = m;
            break;
          }
      if (final_n < 0 && exact) {
        final_n = good;
      } else if (final_n < 0 && n >= max_iter) {
        final_n = n;
        error_code = 1;  // iteration.h:80
      }
      if (final_n >= 0) break;
      checked = n;
      if (nev == 1 && !params->fixed_iterations && n >= 2) {
        double lam = 0.0;
        const double err = lowest_ritz_error(a.data(), b.data(), n, &lam);
        const double tol = stop_tolerance(fp32) * std::max(1.0, std::fabs(lam));
        predicted = -1;
        if (err_n > 0 && err > 0.0 && err_prev > err && err > tol) {
          const double rate = (std::log(err) - std::log(err_prev)) / (n - err_n);  // < 0
          predicted = n + (int)std::ceil((std::log(tol) - std::log(err)) / rate);
        }
        err_prev = err;
        err_n = n;
      }
      if (pending.empty()) queue_to(next_target(queued));
    }
  }
  TValues t = stop_analysis(a, b, final_n, fp32, nev);
  std::vector<double> ev;
  if (!fp32) {
    ev.assign(t.ev.begin(), t.ev.begin() + std::min<size_t>(nev, t.ev.size()));
    if (want_vectors)
      out.vectors = ritz_vectors(lz.get(), a, b, final_n, std::vector<int>(t.index.begin(), t.index.begin() + ev.size()),
                                 t.all);
  } else {
    // Rayleigh quotients of the Ritz vectors; values that agree to 1e-9
    // relative are the same level reached twice and are reported once.
    size_t next = 0;
    std::vector<std::pair<double, std::unique_ptr<alps_ed_vector_s>>> kept;
    while ((int)kept.size() < nev) {
      // more candidates than the stopping test listed: the full analysis
      // (whose first entries are the same) lists them all
      if (next >= t.ev.size() && !t.complete) t = analyze_prefix(a, b, final_n, fp32);
      if (next >= t.ev.size()) break;
      const size_t cnt = std::min<size_t>(t.ev.size() - next, (size_t)(nev - (int)kept.size()));
      std::vector<int> idx(t.index.begin() + next, t.index.begin() + next + cnt);
      std::vector<std::unique_ptr<alps_ed_vector_s>> vs;
      std::vector<double> r = refine_ritz(lz.get(), a, b, final_n, idx, t.all, want_vectors ? &vs : nullptr);
      next += cnt;
      for (size_t i = 0; i < r.size(); ++i) {
        bool dup = false;
        for (const auto& e : kept)
          if (std::fabs(r[i] - e.first) <= 1e-9 * std::max(1.0, std::fabs(e.first))) dup = true;
        if (!dup) kept.emplace_back(r[i], want_vectors ? std::move(vs[i]) : nullptr);
      }
      std::sort(kept.begin(), kept.end(),
                [](const std::pair<double, std::unique_ptr<alps_ed_vector_s>>& x,
                   const std::pair<double, std::unique_ptr<alps_ed_vector_s>>& y) { return x.first < y.first; });
    }
    for (auto& e : kept) {
      ev.push_back(e.first);
      if (want_vectors) out.vectors.push_back(std::move(e.second));
    }
  }
  result->n_eigenvalues = (int)ev.size();
  for (size_t i = 0; i < ev.size(); ++i) result->eigenvalues[i] = ev[i];
  result->iterations = final_n;
  result->error_code = error_code;
  result->storage = fp32 ? ALPS_ED_F32 : ALPS_ED_F64;
  result->reorthogonalizations = reorthogonalized_steps(lz.get(), final_n - 1);
  result->seconds = seconds();
  return out;
}

}  // namespace

extern "C" int alps_ed_sparsediag(const alps_ed_model* model, const alps_ed_solve_params* params,
                                  alps_ed_solve_result* result) {
  return guarded([&] {
    require(params && result, "null argument");
    SectorSolve s = solve_sector(model, params, false);
    *result = s.r;
  });
}

extern "C" int alps_ed_sparsediag_momentum(const alps_ed_model* model, int n_translations, const int* perm,
                                           const double* phase_re, const double* phase_im,
                                           const alps_ed_solve_params* params, alps_ed_solve_result* result) {
  return guarded([&] {
    require(model && params && result, "null argument");
    const Translations t = read_translations(model->n_sites, n_translations, perm, phase_re, phase_im);
    const SectorLayout L = build_layout(model_spec(model));
    const BlochOrbits o = bloch_orbits(L, t);
    const MomentumInput mi{&L, &o, &t};
    SectorSolve s = solve_sector(model, params, false, &mi);
    *result = s.r;
  });
}

namespace {

std::string json_list(const std::vector<double>& v) {
  std::string o = "[";
  for (size_t i = 0; i < v.size(); ++i) o += (i ? ", " : "") + json_number(v[i]);
  return o + "]";
}

std::string json_params(const Parameters& p) {
  std::string js = "{";
  bool first = true;
  for (const auto& kv : p) {
    js += (first ? "" : ", ") + json_string(kv.first) + ": " + json_string(kv.second);
    first = false;
  }
  return js + "}";
}

// Device sums of one eigenvector for the measurements.
SpinSums spin_sums(alps_ed_operator op, alps_ed_vector v, const std::vector<std::pair<int, int>>& flips) {
  SpinSums s;
  const int N = op->layout.N;
  s.sz.resize(N);
  s.szsz.resize((size_t)N * N);
  check_status(alps_ed_vector_spin_moments(op, v, &s.n2, s.sz.data(), s.szsz.data()));
  if (!flips.empty()) {
    std::vector<int> a, b;
    for (const auto& f : flips) {
      a.push_back(f.first);
      b.push_back(f.second);
    }
    std::vector<double> f(flips.size());
    check_status(alps_ed_vector_flip_pairs(op, v, (int)flips.size(), a.data(), b.data(), f.data()));
    for (size_t i = 0; i < flips.size(); ++i) s.flip[flips[i]] = f[i];
  }
  return s;
}

// What a task measures, resolved on the host: operator pairs, labels and
// the unsupported requests.
std::string json_measure_plan(const MeasureSpec& spec, const LatticeGraph& g) {
  const MeasurePlan plan = measure_plan(spec, g);
  auto names = [](const std::map<std::string, std::string>& m) {
    std::string o = "{";
    for (const auto& kv : m) o += (o.size() > 1 ? ", " : "") + json_string(kv.first) + ": " + json_string(kv.second);
    return o + "}";
  };
  auto pairs = [](const std::map<std::string, std::pair<std::string, std::string>>& m) {
    std::string o = "{";
    for (const auto& kv : m)
      o += (o.size() > 1 ? ", " : "") + json_string(kv.first) + ": [" + json_string(kv.second.first) + ", " +
           json_string(kv.second.second) + "]";
    return o + "}";
  };
  std::vector<std::string> dl, kl;
  for (const auto& pr : distance_pairs(g))
    dl.push_back(pr.first < 0 ? std::string()
                              : coordinate_string(site_coordinate(g, pr.first)) + " -- " +
                                    coordinate_string(site_coordinate(g, pr.second)));
  for (const auto& k : momenta(g)) kl.push_back(coordinate_string(k));
  std::string un = "[";
  for (size_t i = 0; i < plan.unsupported.size(); ++i) un += (i ? ", " : "") + plan.unsupported[i];
  return "{\"average\": " + names(spec.average) + ", \"local\": " + names(spec.local) +
         ", \"correlations\": " + pairs(spec.correlation) + ", \"structure_factor\": " +
         pairs(spec.structure_factor) + ", \"flip_pairs\": " + std::to_string(plan.flips.size()) +
         ", \"distance_labels\": " + json_labels(dl) + ", \"momentum_labels\": " + json_labels(kl) +
         ", \"unsupported\": " + un + "]}";
}

void emit(const std::string& js, char* out, size_t out_len, size_t* needed) {
  if (needed) *needed = js.size() + 1;
  if (out && out_len) {
    const size_t n = std::min(out_len - 1, js.size());
    std::memcpy(out, js.data(), n);
    out[n] = 0;
  }
}

}  // namespace

namespace {

// TRANSLATION_SYMMETRY as a boolean parameter (default true,
// applications/diag/diag.h:91 and factory.C:25).
// A boolean parameter (true/false/yes/no or an expression), `def` when
// absent.
bool bool_param(const Parameters& p, const Evaluator& ev, const char* name, bool def) {
  if (!p.count(name)) return def;
  std::string v = p.at(name);
  for (char& c : v) c = (char)std::tolower((unsigned char)c);
  if (v == "true" || v == "yes") return true;
  if (v == "false" || v == "no") return false;
  return ev.value(v) != 0.0;
}

bool translation_symmetry(const Parameters& p, const Evaluator& ev) {
  return bool_param(p, ev, "TRANSLATION_SYMMETRY", true);
}

// The momenta a task loops over (DiagMatrix::dostep, diag.h:90-97,117-120):
// every momentum of the lattice's periodic directions unless
// TRANSLATION_SYMMETRY is false, or the one TOTAL_MOMENTUM given (then
// `loop` is false); none on lattices without periodic directions.
std::vector<std::vector<double>> task_momenta(const Parameters& p, const Evaluator& ev, const LatticeGraph& g,
                                              bool& loop) {
  loop = false;
  std::vector<std::vector<double>> ks;
  if (ev.defined("TOTAL_MOMENTUM")) {
    if (g.extent.empty()) throw std::invalid_argument("TOTAL_MOMENTUM needs a LATTICE, not a GRAPH");
    std::vector<double> k;
    std::istringstream is(p.at("TOTAL_MOMENTUM"));
    for (std::string tok; is >> tok;) k.push_back(ev.value(tok));
    ks.push_back(k);
  } else if (translation_symmetry(p, ev)) {
    ks = translation_momenta(g);
    loop = !ks.empty();
  }
  return ks;
}

// alps::write_vector with its default precision of 20 digits
// (src/alps/utility/vectorio.hpp:170-180): the TOTAL_MOMENTUM value of a
// momentum sector.
std::string momentum_string(const std::vector<double>& k) {
  std::string o;
  for (size_t i = 0; i < k.size(); ++i) {
    char b[48];
    std::snprintf(b, sizeof b, "%.20g", k[i]);
    o += (i ? " " : "") + std::string(b);
  }
  return o;
}

}  // namespace

extern "C" int alps_ed_describe_parameter_file(const char* text, char* out, size_t out_len, size_t* needed) {
  return guarded([&] {
    require(text != nullptr, "parameter text is null");
    const std::vector<Parameters> tasks = parse_input(text);
    std::string js = "{\"tasks\": [";
    for (size_t ti = 0; ti < tasks.size(); ++ti) {
      const Parameters& p = tasks[ti];
      js += ti ? ", {" : "{";
      js += "\"parameters\": " + json_params(p);
      if (p.count("LATTICE") || p.count("GRAPH")) {
        const TaskModel m = task_model(p);
        std::vector<double> s1(m.s1.begin(), m.s1.end()), s2(m.s2.begin(), m.s2.end()), up;
        const TaskSectors ts = task_sectors(p, m.n_sites);
        for (int u : ts.n_up) up.push_back(u);
        js += ", \"n_sites\": " + std::to_string(m.n_sites) + ", \"bond_s\": " + json_list(s1) +
              ", \"bond_t\": " + json_list(s2) + ", \"jz\": " + json_list(m.jz) + ", \"jxy\": " + json_list(m.jxy) +
              ", \"k\": " + json_list(m.k) + ", \"h\": " + json_list(m.h) + ", \"d\": " + json_list(m.d) +
              ", \"n_up\": " + json_list(up) + ", \"whole_space\": " + (ts.whole_space ? "true" : "false");
        // vertex types and coordinates (lattices only), as the lattice
        // library's graphs give them (src/alps/lattice/graph.h:64-108)
        std::vector<double> types(m.graph.site_type.begin(), m.graph.site_type.end()), bt;
        for (const LatticeBond& b : m.graph.bonds) bt.push_back(b.type);
        js += ", \"site_type\": " + json_list(types) + ", \"bond_type\": " + json_list(bt);
        if (!m.graph.extent.empty()) {
          js += ", \"coordinates\": [";
          for (int s = 0; s < m.n_sites; ++s) js += (s ? ", " : "") + json_list(site_coordinate(m.graph, s));
          js += "]";
        }
        bool loop = false;
        const std::vector<std::vector<double>> ks = task_momenta(p, Evaluator(p), m.graph, loop);
        if (!ks.empty()) {
          js += ", \"momenta\": [";
          for (size_t q = 0; q < ks.size(); ++q) js += (q ? ", " : "") + json_list(ks[q]);
          js += "]";
        }
        const MeasureSpec ms = measure_spec(p);
        require_geometry(ms, m.graph);
        if (ms.any()) js += ", \"measure\": " + json_measure_plan(ms, m.graph);
      }
      js += "}";
    }
    js += "]}";
    emit(js, out, out_len, needed);
  });
}


extern "C" int alps_ed_run_parameter_file(const char* text, char* out, size_t out_len, size_t* needed) {
  return guarded([&] {
    require(text != nullptr, "parameter text is null");
    const std::vector<Parameters> tasks = parse_input(text);
    std::string js = "{\"tasks\": [";
    for (size_t ti = 0; ti < tasks.size(); ++ti) {
      const Parameters& p = tasks[ti];
      const TaskModel m = task_model(p);
      Evaluator ev(p);
      alps_ed_solve_params sp{};
      sp.max_iter = ev.defined("MAX_ITERATIONS") ? (int)std::lround(ev.param("MAX_ITERATIONS")) : 0;
      sp.nev = ev.defined("NUMBER_EIGENVALUES") ? (int)std::lround(ev.param("NUMBER_EIGENVALUES")) : 1;
      sp.fixed_iterations = 0;
      sp.seed = ev.defined("ED_SEED") ? (uint64_t)std::llround(ev.param("ED_SEED")) : 0u;
      sp.dtype = (p.count("ED_PRECISION") && p.at("ED_PRECISION") == "f32") ? ALPS_ED_F32 : ALPS_ED_F64;
      js += ti ? ", {" : "{";
      js += "\"parameters\": " + json_params(p) + ", \"n_sites\": " + std::to_string(m.n_sites) + ", \"sectors\": [";
      const TaskSectors ts = task_sectors(p, m.n_sites);
      // momenta (DiagMatrix::dostep, diag.h:90-97,117-120): every momentum of
      // the lattice's periodic directions unless TRANSLATION_SYMMETRY is
      // false, or the one TOTAL_MOMENTUM given; the loop runs inside the Sz
      // loop.  Lattices without periodic directions have none.
      bool loop_momenta = false;
      const std::vector<std::vector<double>> ks = task_momenta(p, ev, m.graph, loop_momenta);
      const bool momentum = !ks.empty();
      std::vector<alps_ed_model> models;
      for (int n_up : ts.n_up)
        models.push_back(alps_ed_model{m.n_sites, n_up, (int)m.s1.size(), m.s1.data(), m.s2.data(), m.jz.data(),
                                       m.jxy.data(), m.k.data(), m.h.data(), m.d.data()});
      const size_t nk = momentum ? ks.size() : 1;
      // one result per (Sz sector, momentum), the momentum fastest
      std::vector<alps_ed_solve_result> res(models.size() * nk);
      // Eigenvector measurements (DiagMatrix::perform_measurements,
      // diag.h:191-314) need the eigenvectors of every sector: those tasks
      // solve their sectors one by one with the Lanczos basis kept.  In a
      // momentum sector the eigenvectors are expanded into the fixed-Sz
      // configuration basis and measured there; local values are not
      // reported there, as in the reference (diag.h:212).
      const MeasureSpec ms = measure_spec(p);
      require_geometry(ms, m.graph);
      const bool measure = ms.any() && !ts.whole_space;
      std::vector<std::string> meas_json(res.size());
      std::vector<MeasureValues> meas_val(res.size());
      if (momentum && measure) {
        std::vector<Translations> trans;
        for (const auto& k : ks) trans.push_back(lattice_translations(m.graph, k));
        const MeasurePlan plan = measure_plan(ms, m.graph);
        for (size_t i = 0; i < models.size(); ++i) {
          const SectorLayout L = build_layout(model_spec(&models[i]));
          const BlochOrbits orb = bloch_orbits(L, trans[0]);
          std::unique_ptr<alps_ed_operator_s> plain;  // the fixed-Sz sector, for the spin sums
          for (size_t q = 0; q < nk; ++q) {
            const MomentumInput mi{&L, &orb, &trans[q]};
            SectorSolve sv = solve_sector(&models[i], &sp, true, &mi);
            res[i * nk + q] = sv.r;
            if (sv.r.dim == 0) continue;
            if (!plain) {
              alps_ed_operator h = nullptr;
              check_status(alps_ed_operator_create(&models[i], &h));
              plain.reset(h);
            }
            std::vector<SpinSums> sums;
            for (auto& v : sv.vectors) {
              alps_ed_vector re = nullptr, im = nullptr;
              check_status(alps_ed_vector_create(L.D, ALPS_ED_F64, &re));
              std::unique_ptr<alps_ed_vector_s> hre(re);
              check_status(alps_ed_vector_create(L.D, ALPS_ED_F64, &im));
              std::unique_ptr<alps_ed_vector_s> him(im);
              check_status(alps_ed_vector_bloch_expand(sv.op.get(), v.get(), re, im));
              // <v|O|v> of a complex v: the sums of its real and imaginary
              // parts (the real part of each flip amplitude)
              SpinSums a = spin_sums(plain.get(), re, plan.flips), b = spin_sums(plain.get(), im, plan.flips);
              a.n2 += b.n2;
              for (size_t x = 0; x < a.sz.size(); ++x) a.sz[x] += b.sz[x];
              for (size_t x = 0; x < a.szsz.size(); ++x) a.szsz[x] += b.szsz[x];
              for (auto& f : a.flip) f.second += b.flip[f.first];
              sums.push_back(std::move(a));
            }
            meas_val[i * nk + q] = evaluate_measurements(ms, m.graph, sums);
            meas_val[i * nk + q].local.clear();
            meas_json[i * nk + q] = measurements_json(meas_val[i * nk + q], plan);
          }
        }
      } else if (momentum) {
        std::vector<Translations> trans;
        for (const auto& k : ks) trans.push_back(lattice_translations(m.graph, k));
        // Every (Sz, k) sector as one batch when the momentum sectors' bases
        // are small (about 13 bytes per state of each Sz sector and momentum
        // on the host while sizing) and the whole batch fits the free device
        // memory, else one after the other.
        double basis_bytes = 0;
        for (const alps_ed_model& am : models) basis_bytes += 13.0 * (double)binom(am.n_sites, am.n_up) * (double)nk;
        std::vector<std::vector<int>> perm(nk);
        std::vector<std::vector<double>> re(nk), im(nk);
        for (size_t q = 0; q < nk; ++q)
          for (size_t t = 0; t < trans[q].perm.size(); ++t) {
            perm[q].insert(perm[q].end(), trans[q].perm[t].begin(), trans[q].perm[t].end());
            re[q].push_back(trans[q].phase[t].real());
            im[q].push_back(trans[q].phase[t].imag());
          }
        // Sizing builds each Sz sector's orbits (kept for the batch): the
        // Sz sectors are sized in parallel threads.
        uint64_t need = 0;
        if (res.size() > 1 && basis_bytes < 4e9) {
          std::vector<uint64_t> fp(res.size(), 0);
          std::vector<int> st(res.size(), ALPS_ED_OK);
          std::atomic<size_t> next{0};
          auto work = [&] {
            for (size_t i = next++; i < models.size(); i = next++)
              for (size_t q = 0; q < nk; ++q)
                st[i * nk + q] = alps_ed_batch_footprint_momentum(&models[i], (int)trans[q].perm.size(),
                                                                  perm[q].data(), re[q].data(), im[q].data(), &sp,
                                                                  sp.dtype, &fp[i * nk + q]);
          };
          std::vector<std::thread> th;
          const size_t nth = std::min<size_t>(models.size(), 8);
          for (size_t t = 1; t < nth; ++t) th.emplace_back(work);
          work();
          for (auto& t : th) t.join();
          for (size_t i = 0; i < res.size(); ++i) {
            if (st[i] != ALPS_ED_OK)
              check_status(alps_ed_batch_footprint_momentum(&models[i / nk], (int)trans[i % nk].perm.size(),
                                                            perm[i % nk].data(), re[i % nk].data(),
                                                            im[i % nk].data(), &sp, sp.dtype, &fp[i]));
            need += fp[i] + (1u << 20);
          }
        }
        size_t free_b = 0, total_b = 0;
        hip_check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
        if (need > 0 && need < 0.8 * (double)free_b) {
          alps_ed_batch_config cfg{sp.dtype, need, 0, 8, 0, 0, 0};
          alps_ed_batch bh = nullptr;
          check_status(alps_ed_batch_create(&cfg, &bh));
          std::unique_ptr<alps_ed_batch_s, int (*)(alps_ed_batch)> hold(bh, alps_ed_batch_destroy);
          std::vector<uint64_t> tickets(res.size());
          for (size_t i = 0; i < models.size(); ++i)
            for (size_t q = 0; q < nk; ++q)
              check_status(alps_ed_batch_submit_momentum(bh, &models[i], (int)trans[q].perm.size(), perm[q].data(),
                                                         re[q].data(), im[q].data(), &sp, &tickets[i * nk + q]));
          check_status(alps_ed_batch_run(bh));
          for (size_t i = 0; i < res.size(); ++i) {
            int ready = 0;
            check_status(alps_ed_batch_result(bh, tickets[i], &res[i], &ready));
            if (!ready || res[i].error_code < 0) throw DeviceError("sector solve did not complete");
          }
        } else {
          for (size_t i = 0; i < models.size(); ++i) {
            const SectorLayout L = build_layout(model_spec(&models[i]));
            const BlochOrbits orb = bloch_orbits(L, trans[0]);
            for (size_t q = 0; q < nk; ++q) {
              const MomentumInput mi{&L, &orb, &trans[q]};
              res[i * nk + q] = solve_sector(&models[i], &sp, false, &mi).r;
            }
          }
        }
      } else if (measure) {
        const MeasurePlan plan = measure_plan(ms, m.graph);
        for (size_t i = 0; i < models.size(); ++i) {
          SectorSolve sv = solve_sector(&models[i], &sp, true);
          res[i] = sv.r;
          if (sv.r.dim == 0) continue;
          std::vector<SpinSums> sums;
          for (auto& v : sv.vectors) sums.push_back(spin_sums(sv.op.get(), v.get(), plan.flips));
          meas_val[i] = evaluate_measurements(ms, m.graph, sums);
          meas_json[i] = measurements_json(meas_val[i], plan);
        }
      } else {
        // All sectors of the task are solved together as one batch
        // (DiagMatrix::dostep's sector loop, diag.h:86-138) when their
        // footprint fits the free device memory, else one after the other.
        uint64_t need = 0;
        for (const alps_ed_model& am : models) {
          uint64_t f = 0;
          check_status(alps_ed_batch_footprint(&am, &sp, sp.dtype, &f));
          need += f + (1u << 20);
        }
        size_t free_b = 0, total_b = 0;
        hip_check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
        if (models.size() > 1 && need < 0.8 * (double)free_b) {
          alps_ed_batch_config cfg{sp.dtype, need, 0, 8, 0, 0, 0};
          alps_ed_batch bh = nullptr;
          check_status(alps_ed_batch_create(&cfg, &bh));
          std::unique_ptr<alps_ed_batch_s, int (*)(alps_ed_batch)> hold(bh, alps_ed_batch_destroy);
          std::vector<uint64_t> tickets(models.size());
          for (size_t i = 0; i < models.size(); ++i)
            check_status(alps_ed_batch_submit(bh, &models[i], &sp, &tickets[i]));
          check_status(alps_ed_batch_run(bh));
          for (size_t i = 0; i < models.size(); ++i) {
            int ready = 0;
            check_status(alps_ed_batch_result(bh, tickets[i], &res[i], &ready));
            if (!ready || res[i].error_code < 0) throw DeviceError("sector solve did not complete");
          }
        } else {
          for (size_t i = 0; i < models.size(); ++i) check_status(alps_ed_sparsediag(&models[i], &sp, &res[i]));
        }
      }
      // sectors of a quantum-number loop carry their Sz (diag.h:100-105), a
      // momentum loop its TOTAL_MOMENTUM (:117-120)
      const bool qn_loop = !ts.whole_space && !ev.defined("Sz_total");
      std::vector<std::vector<double>> evs;
      std::vector<std::string> entries;
      std::vector<XmlSector> xsec;
      // per momentum, for the whole-space merge
      std::vector<std::vector<std::vector<double>>> by_k(nk);
      std::vector<uint64_t> dim_k(nk, 0);
      std::vector<int> it_k(nk, 0), err_k(nk, 0);
      std::vector<double> sec_k(nk, 0.0);
      for (size_t i = 0; i < res.size(); ++i) {
        const alps_ed_solve_result& r = res[i];
        const int n_up = ts.n_up[i / nk];
        const size_t q = i % nk;
        if (r.dim == 0) continue;  // empty subspaces are skipped (diag.h:115)
        std::vector<double> e(r.eigenvalues, r.eigenvalues + r.n_eigenvalues);
        evs.push_back(e);
        by_k[q].push_back(e);
        dim_k[q] += r.dim;
        it_k[q] += r.iterations;
        err_k[q] = std::max(err_k[q], r.error_code);
        sec_k[q] += r.seconds;
        entries.push_back("\"Sz_total\": " + json_number(n_up - 0.5 * m.n_sites) + ", \"n_up\": " +
                          std::to_string(n_up) + (momentum ? ", \"momentum\": " + json_list(ks[q]) : std::string()) +
                          ", \"dim\": " + std::to_string(r.dim) + ", \"iterations\": " +
                          std::to_string(r.iterations) + ", \"error_code\": " + std::to_string(r.error_code) +
                          ", \"seconds\": " + json_number(r.seconds) + ", \"eigenvalues\": " + json_list(e) +
                          (measure ? ", \"measurements\": " + meas_json[i] : std::string()));
        XmlSector xs;
        if (qn_loop) xs.qn.push_back({"Sz", half_integer_string(2 * n_up - m.n_sites)});
        if (loop_momenta) xs.qn.push_back({"TOTAL_MOMENTUM", momentum_string(ks[q])});
        xs.ev = e;
        if (measure) xs.meas = &meas_val[i];
        xsec.push_back(xs);
      }
      if (ts.whole_space) {
        // one subspace without Sz (per momentum): the merged spectrum
        evs.clear();
        entries.clear();
        xsec.clear();
        for (size_t q = 0; q < nk; ++q) {
          if (by_k[q].empty()) continue;
          const std::vector<double> e = merge_spectra(by_k[q], sp.nev, 1e-9);
          evs.push_back(e);
          XmlSector xs;
          if (loop_momenta) xs.qn.push_back({"TOTAL_MOMENTUM", momentum_string(ks[q])});
          xs.ev = e;
          xsec.push_back(xs);
          entries.push_back("\"Sz_total\": null, \"n_up\": null" +
                            (momentum ? ", \"momentum\": " + json_list(ks[q]) : std::string()) +
                            ", \"dim\": " + std::to_string(dim_k[q]) + ", \"iterations\": " +
                            std::to_string(it_k[q]) + ", \"error_code\": " + std::to_string(err_k[q]) +
                            ", \"seconds\": " + json_number(sec_k[q]) + ", \"eigenvalues\": " + json_list(e));
        }
      }
      for (size_t i = 0; i < entries.size(); ++i) js += (i ? ", {" : "{") + entries[i] + "}";
      double g = 0, gap = 0;
      ground_and_gap(evs, g, gap);
      js += "], \"ground_state_energy\": " + json_number(g) + ", \"energy_gap\": " + json_number(gap);
      if (ms.any() && !measure)
        js += std::string(", \"measurement_error\": \"eigenvector measurements need Sz-conserved sectors") +
              " (CONSERVED_QUANTUMNUMBERS=\\\"Sz\\\" or Sz_total)\"";
      const int nev_xml = ev.defined("NUMBER_EIGENVALUES") ? sp.nev : 0;
      // <EIGENSTATES> when averages are measured or MEASURE_ENERGY (default
      // true) holds (DiagTask::write_xml_body, diag.hpp:141)
      const bool states = ms.any() || bool_param(p, ev, "MEASURE_ENERGY", true);
      js += ", \"xml\": " +
            json_string(task_xml(p, xsec, nev_xml, !std::isnan(g), g, !std::isnan(gap), gap, states));
      js += "}";
    }
    js += "]}";
    emit(js, out, out_len, needed);
  });
}
