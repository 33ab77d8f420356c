// ed_batch.hip -- the batched solver and request loop behind the
// alps_ed_batch_* C ABI (include/alps_ed.h).
//
// A batch holds requests (one do_subspace each: a sector of a spin-1/2
// Hamiltonian) in three stages:
//   pending  -- submitted, waiting for HBM;
//   active   -- admitted: its tables, Lanczos vectors, scalars and partials
//               live in one region of the batch's HBM arena, sized when it
//               is admitted (admission is by footprint, first come first
//               served);
//   done     -- result recorded, region freed.
// alps_ed_batch_progress runs one round: admit what fits, start the new
// problems, advance every running problem by `steps_per_round` Lanczos
// steps (four batched launches per step, no host round trip), then test
// each problem's stopping rule on the host (in parallel threads) from the
// alpha/beta the kernels mirror into mapped host memory, refine and retire
// the finished ones.  A problem's arithmetic never depends on the rest of
// the batch, so its result is the same alone and in any mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "alps_ed.h"
#include "ed_batch.hpp"
#include "ed_bloch.hpp"
#include "ed_host.hpp"
#include "ed_kernels.hpp"

using namespace alps_ed;
using namespace alps_ed::host;

namespace {

constexpr size_t kAlign = 256;
inline size_t align_up(size_t x) { return (x + kAlign - 1) & ~(kAlign - 1); }

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// First-fit sub-allocator of byte offsets in one region.
class Arena {
 public:
  void init(size_t size) {
    size_ = size;
    used_ = 0;
    free_.clear();
    if (size) free_[0] = size;
  }
  bool alloc(size_t bytes, size_t* off) {
    bytes = align_up(bytes);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < bytes) continue;
      *off = it->first;
      const size_t rest = it->second - bytes, next = it->first + bytes;
      free_.erase(it);
      if (rest) free_[next] = rest;
      used_ += bytes;
      return true;
    }
    return false;
  }
  void release(size_t off, size_t bytes) {
    bytes = align_up(bytes);
    used_ -= bytes;
    auto it = free_.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
  size_t used() const { return used_; }
  size_t size() const { return size_; }

 private:
  size_t size_ = 0, used_ = 0;
  std::map<size_t, size_t> free_;
};

// Persistent worker threads for the per-problem host analysis.
class Pool {
 public:
  explicit Pool(int n) {
    for (int i = 0; i < n; ++i) th_.emplace_back([this] { loop(); });
  }
  ~Pool() {
    {
      std::lock_guard<std::mutex> g(m_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto& t : th_) t.join();
  }
  // fn(i) for i in [0, n), on the workers and the calling thread
  void parallel_for(size_t n, const std::function<void(size_t)>& fn) {
    if (n == 0) return;
    if (th_.empty() || n == 1) {
      for (size_t i = 0; i < n; ++i) fn(i);
      return;
    }
    {
      std::lock_guard<std::mutex> g(m_);
      fn_ = &fn;
      n_ = n;
      next_ = 0;
      busy_ = (int)th_.size();
      ++gen_;
    }
    cv_.notify_all();
    run_items();
    std::unique_lock<std::mutex> lk(m_);
    done_cv_.wait(lk, [this] { return busy_ == 0; });
    fn_ = nullptr;
  }

 private:
  void run_items() {
    for (;;) {
      const size_t i = next_.fetch_add(1);
      if (i >= n_) break;
      (*fn_)(i);
    }
  }
  void loop() {
    uint64_t seen = 0;
    for (;;) {
      {
        std::unique_lock<std::mutex> lk(m_);
        cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
        if (stop_) return;
        seen = gen_;
      }
      run_items();
      {
        std::lock_guard<std::mutex> g(m_);
        if (--busy_ == 0) done_cv_.notify_all();
      }
    }
  }
  std::vector<std::thread> th_;
  std::mutex m_;
  std::condition_variable cv_, done_cv_;
  const std::function<void(size_t)>* fn_ = nullptr;
  size_t n_ = 0;
  std::atomic<size_t> next_{0};
  int busy_ = 0;
  uint64_t gen_ = 0;
  bool stop_ = false;
};

// Device copy of the tables that depend only on the low-word size and the
// bond topology (low-word ranks and states, class offsets, flip tables).
struct SharedTables {
  DevBuf buf;
  std::shared_ptr<const FlipTables> keep;  // keeps the key alive
  const uint32_t *lo_rank = nullptr, *lo_state = nullptr, *lo_class_off = nullptr, *nbp_off = nullptr,
                 *xn_off = nullptr;
  const uint16_t *nbp16 = nullptr, *xn16 = nullptr;
};

enum State { kPending, kRunning, kRefining, kDone };

struct Problem {
  uint64_t ticket = 0;
  ModelSpec spec;
  SectorLayout layout;
  std::shared_ptr<const Translations> trans;  // set for a momentum sector
  BlochSector bloch;                          // its basis (after sizing)
  uint64_t D = 0;                             // dimension (representatives in a momentum sector)
  uint64_t n = 0;                             // Lanczos vector length: D, or 2 D realified
  uint32_t n_hv = 0;                          // H.v work items per step
  bool fp32 = false;  // storage lane
  int max_iter = 0, nev = 1;
  bool fixed = false;
  uint64_t seed = 0;
  State state = kPending;
  // HBM region and mapped host region
  size_t dev_off = 0, dev_bytes = 0, host_off = 0, host_bytes = 0;
  BatchJob job{};
  uint32_t n_chunks = 0, cap = 0;
  int nvr = 0;                  // Ritz vectors per refinement pass
  double* coef = nullptr;       // device [nvr * (max_iter+1)]
  double* ritz = nullptr;       // device [nvr * D]
  double* hy = nullptr;         // device [D]
  double* ritz_part = nullptr;  // mapped [nvr * n_chunks]
  double* hv_part = nullptr;    // mapped [n_seg]
  double* coef_host = nullptr;  // mapped staging of coef
  double* host_alpha = nullptr;
  double* host_beta = nullptr;
  // recurrence and stopping rule
  int j = 0;           // alpha values produced
  double lam_prev = std::numeric_limits<double>::quiet_NaN();  // lowest Ritz value at the last test
  double lam_drop = 0;  // its decrease between the last two tests
  int tested_j = 0;     // value of j at the last stopping test
  int final_n = -1;
  int error_code = 0;
  TValues tv;
  int tv_n = -1;            // length of T that tv analyses
  std::vector<double> ev;   // reported eigenvalues
  size_t next_cand = 0;     // next candidate of tv to refine
  std::vector<int> cand;    // tv positions in the current refinement pass
  double t_submit = 0, t_admit = 0, t_done = 0;
  bool sized = false;  // layout built
  alps_ed_solve_result result{};
};

}  // namespace

struct alps_ed_batch_s {
  alps_ed_batch_config cfg{};
  hipStream_t stream = nullptr;
  void* dev_base = nullptr;
  char* host_base = nullptr;   // mapped, pinned
  char* host_dev = nullptr;    // device view of host_base
  Arena arena, harena;
  std::unique_ptr<Pool> pool;
  std::deque<std::unique_ptr<Problem>> pending;
  std::vector<std::unique_ptr<Problem>> active;
  std::unordered_map<uint64_t, alps_ed_solve_result> results;
  std::vector<uint64_t> completed;
  uint64_t next_ticket = 1;
  uint64_t rounds = 0, steps = 0, solves_done = 0;
  uint64_t hv_rows[2] = {0, 0};  // rows of Lanczos-step H.v, FP64 / FP32 storage
  // the round's steps as a HIP graph (cfg.graphs): the launch key of the
  // last round and of the captured graph
  std::vector<uint64_t> last_key, graph_key;
  hipGraphExec_t graph_exec = nullptr;
  uint64_t graph_rounds = 0, graph_captures = 0;
  double t_admit = 0, t_steps = 0, t_decide = 0, t_refine = 0;
  // device work lists (rebuilt every round)
  DevBuf jobs_d, items_d, list_d;
  size_t jobs_cap = 0, items_cap = 0, list_cap = 0;
  // pinned staging for the work lists and the tables of new problems
  void* stage = nullptr;
  size_t stage_cap = 0;
  void* tstage = nullptr;
  size_t tstage_cap = 0;
  DevBuf tdev;  // the device copy of tstage the scatter launch reads
  size_t tdev_cap = 0;
  void* rstage = nullptr;  // pinned stage of a refinement pass's uploads
  size_t rstage_cap = 0;
  DevBuf rdev;  // its device copy
  size_t rdev_cap = 0;
  // shared low-word and flip tables, one device copy per lattice topology
  std::map<std::pair<const void*, const void*>, std::unique_ptr<SharedTables>> shared;

  ~alps_ed_batch_s() {
    if (stream) (void)hipStreamSynchronize(stream);
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (dev_base) (void)hipFree(dev_base);
    if (host_base) (void)hipHostFree(host_base);
    if (stage) (void)hipHostFree(stage);
    if (tstage) (void)hipHostFree(tstage);
    if (rstage) (void)hipHostFree(rstage);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

// Bytes of the per-problem layout tables (the low-word and flip tables are
// shared by all problems of a lattice and uploaded once, see SharedTables).
size_t table_bytes(const SectorLayout& L) {
  size_t s = 0;
  auto add = [&](size_t n, size_t es) { s += align_up(std::max<size_t>(n, 1) * es); };
  add(L.hi_base.size(), 4);
  add(L.seg_hi.size(), 4);
  add(L.seg_start.size(), 4);
  add(L.proc_hi.size(), 4);
  add(L.diag_lo_cls.size(), 8);
  add(L.diag_hi.size(), 8);
  add(L.ll_t.size(), 8);
  add(L.hh_off.size(), 4);
  add(L.hh_ent.size(), sizeof(HHEntry));
  add(L.xb_c.size(), sizeof(XBondConst));
  add(L.xb_nb.size(), 4);
  add(L.site_dn.size(), 8);
  add(L.site_up.size(), 8);
  add(L.bond_p1.size(), 4);
  add(L.bond_p2.size(), 4);
  add(L.bond_par.size(), 8);
  add(L.bond_anti.size(), 8);
  add(L.bond_t.size(), 8);
  return s;
}


// Bytes of a momentum sector's basis arrays (none for a fixed-Sz sector).
size_t bloch_bytes(const BlochSector& b) {
  if (b.rep_state.empty()) return 0;
  return align_up(b.full_pk.size() * 8) + align_up(b.norm_tab.size() * 8) + align_up(b.rep_state.size() * 4) +
         align_up(b.norm.size() * 8) + 2 * align_up(b.ph_re.size() * 8);
}

size_t problem_table_bytes(const Problem& p) { return table_bytes(p.layout) + bloch_bytes(p.bloch); }

// Lays the tables out at [base, base + table_bytes) of the device, writing
// their bytes to the same offsets of the host staging area `stage`, and
// returns the device layout; one copy of the staging area uploads them.
DevLayout place_tables(const SectorLayout& L, char* base, char* stage, const SharedTables& sh) {
  DevLayout d{};
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
  d.lo_rank = sh.lo_rank;
  d.lo_state = sh.lo_state;
  d.lo_class_off = sh.lo_class_off;
  d.nbp16 = sh.nbp16;
  d.nbp_off = sh.nbp_off;
  d.xn16 = sh.xn16;
  d.xn_off = sh.xn_off;
  size_t o = 0;
  auto put = [&](const auto& v) {
    using T = typename std::decay_t<decltype(v)>::value_type;
    if (!v.empty()) std::memcpy(stage + o, v.data(), v.size() * sizeof(T));
    const T* dst = reinterpret_cast<const T*>(base + o);
    o += align_up(std::max<size_t>(v.size(), 1) * sizeof(T));
    return dst;
  };
  d.hi_base = put(L.hi_base);
  d.seg_hi = put(L.seg_hi);
  d.seg_start = put(L.seg_start);
  d.proc_hi = put(L.proc_hi);
  d.diag_lo_cls = put(L.diag_lo_cls);
  d.diag_hi = put(L.diag_hi);
  d.ll_t = put(L.ll_t);
  d.hh_off = put(L.hh_off);
  d.hh_ent = put(L.hh_ent);
  d.xb_c = put(L.xb_c);
  d.xb_nb = put(L.xb_nb);
  d.site_dn = put(L.site_dn);
  d.site_up = put(L.site_up);
  d.bond_p1 = put(L.bond_p1);
  d.bond_p2 = put(L.bond_p2);
  d.bond_par = put(L.bond_par);
  d.bond_anti = put(L.bond_anti);
  d.bond_t = put(L.bond_t);
  return d;
}

// H(0,0) of a one-state sector in the reference's summation order
// (hamiltonian_matrix.hpp:44-58, 269-291, 308-345): explicit 0.0, sites in
// order, then bonds in edge order.
double single_state_energy(const ModelSpec& m) {
  const bool up = m.n_up == m.n_sites;  // the only states are all-down / all-up
  double v = 0.0;
  for (const SiteTerm& s : m.sites) v += up ? s.e_up : s.e_dn;
  for (const BondTerm& b : m.bonds) v += b.e_par;
  return v;
}

struct Footprint {
  size_t dev = 0, host = 0;
};

Footprint footprint(const Problem& p) {
  Footprint f;
  const size_t es = p.fp32 ? 4 : 8;
  const uint64_t stride = (p.n + 63) & ~uint64_t(63);
  f.dev += problem_table_bytes(p);
  f.dev += align_up((size_t)p.cap * stride * es);
  f.dev += 3 * align_up((size_t)(p.max_iter + 2) * 8);  // alpha, beta, scale
  f.dev += align_up((size_t)std::max<uint32_t>(p.n_hv, 1) * 8);
  f.dev += align_up((size_t)p.n_chunks * 8);
  if (p.fp32) {
    f.dev += align_up((size_t)p.nvr * (p.max_iter + 1) * 8);  // coef
    f.dev += align_up((size_t)p.nvr * p.n * 8);              // Ritz vectors
    f.dev += align_up((size_t)p.n * 8);                      // H y
    f.host += align_up((size_t)p.nvr * p.n_chunks * 8);
    f.host += align_up((size_t)std::max<uint32_t>(p.n_hv, 1) * 8);
    f.host += align_up((size_t)p.nvr * (p.max_iter + 1) * 8);
  }
  f.host += 2 * align_up((size_t)(p.max_iter + 2) * 8);  // alpha, beta mirrors
  return f;
}

template <class T>
T* dev_ptr(alps_ed_batch b, size_t off) {
  return reinterpret_cast<T*>(static_cast<char*>(b->dev_base) + off);
}

// Places an admitted problem in its regions; its tables are staged at
// `stage` (table_bytes of them) for the caller to upload.
void place(alps_ed_batch b, Problem& p, char* stage, const SharedTables& sh) {
  char* base = static_cast<char*>(b->dev_base) + p.dev_off;
  BatchJob& J = p.job;
  J = BatchJob{};
  J.L = place_tables(p.layout, base, stage, sh);
  if (!p.bloch.rep_state.empty()) {  // the momentum sector's basis after the layout tables
    size_t o = table_bytes(p.layout);
    auto put = [&](const auto& v) {
      using T = typename std::decay_t<decltype(v)>::value_type;
      std::memcpy(stage + o, v.data(), v.size() * sizeof(T));
      const T* dst = reinterpret_cast<const T*>(base + o);
      o += align_up(v.size() * sizeof(T));
      return dst;
    };
    J.Bl.dk = (uint32_t)p.bloch.rep_state.size();
    J.Bl.full_pk = put(p.bloch.full_pk);
    J.Bl.norm_tab = put(p.bloch.norm_tab);
    J.Bl.rep_state = put(p.bloch.rep_state);
    J.Bl.norm = put(p.bloch.norm);
    J.Bl.ph_re = put(p.bloch.ph_re);
    J.Bl.ph_im = put(p.bloch.ph_im);
  }
  J.n = (uint32_t)p.n;
  J.n_hv = p.n_hv;
  char* q = base + problem_table_bytes(p);
  const size_t es = p.fp32 ? 4 : 8;
  const uint64_t stride = (p.n + 63) & ~uint64_t(63);
  J.vec = q;
  J.stride = stride;
  J.cap = p.cap;
  J.n_chunks = p.n_chunks;
  J.j0 = 0;
  J.jmax = p.max_iter;
  J.seed = p.seed;
  q += align_up((size_t)p.cap * stride * es);
  const size_t sc = align_up((size_t)(p.max_iter + 2) * 8);
  J.alpha = reinterpret_cast<double*>(q);
  J.beta = reinterpret_cast<double*>(q + sc);
  J.scale = reinterpret_cast<double*>(q + 2 * sc);
  q += 3 * sc;
  J.part_hv = reinterpret_cast<double*>(q);
  q += align_up((size_t)std::max<uint32_t>(p.n_hv, 1) * 8);
  J.part_up = reinterpret_cast<double*>(q);
  q += align_up((size_t)p.n_chunks * 8);
  char* h = b->host_base + p.host_off;
  const size_t hoff = (size_t)(h - b->host_base);
  p.host_alpha = reinterpret_cast<double*>(h);
  p.host_beta = reinterpret_cast<double*>(h + sc);
  J.host_alpha = reinterpret_cast<double*>(b->host_dev + hoff);
  J.host_beta = reinterpret_cast<double*>(b->host_dev + hoff + sc);
  size_t ho = 2 * sc;
  if (p.fp32) {
    p.coef = reinterpret_cast<double*>(q);
    q += align_up((size_t)p.nvr * (p.max_iter + 1) * 8);
    p.ritz = reinterpret_cast<double*>(q);
    q += align_up((size_t)p.nvr * p.n * 8);
    p.hy = reinterpret_cast<double*>(q);
    q += align_up((size_t)p.n * 8);
    p.ritz_part = reinterpret_cast<double*>(h + ho);
    ho += align_up((size_t)p.nvr * p.n_chunks * 8);
    p.hv_part = reinterpret_cast<double*>(h + ho);
    ho += align_up((size_t)std::max<uint32_t>(p.n_hv, 1) * 8);
    p.coef_host = reinterpret_cast<double*>(h + ho);
    ho += align_up((size_t)p.nvr * (p.max_iter + 1) * 8);
  }
}

// device view of a mapped host pointer
template <class T>
T* mapped_dev(alps_ed_batch b, T* h) {
  return reinterpret_cast<T*>(b->host_dev + (reinterpret_cast<char*>(h) - b->host_base));
}

void finish_problem(alps_ed_batch b, Problem& p) {
  p.state = kDone;
  p.t_done = now_s();
  alps_ed_solve_result& r = p.result;
  r.dim = p.D;
  r.n_eigenvalues = (int)std::min<size_t>(p.ev.size(), 64);
  for (int i = 0; i < r.n_eigenvalues; ++i) r.eigenvalues[i] = p.ev[i];
  r.iterations = p.final_n < 0 ? 0 : p.final_n;
  r.error_code = p.error_code;
  r.seconds = p.t_done - p.t_admit;
  r.storage = p.fp32 ? ALPS_ED_F32 : ALPS_ED_F64;
  r.reorthogonalizations = 0;
  b->results[p.ticket] = r;
  b->completed.push_back(p.ticket);
  ++b->solves_done;
}

// Stopping rule on the alpha/beta produced so far.  The batch tests
// lanczos_iteration_nlowest at the end of each round, i.e. at step counts
// that are multiples of steps_per_round for every problem (each starts at
// step 0 when admitted), and reports the T matrix of the first such count
// at which the rule holds.  A problem's result therefore does not depend on
// when it was admitted or on the rest of the batch.  A cheap O(m) test of
// the lowest value gates the full T-matrix analysis.
void decide(Problem& p) {
  const int n = p.j;
  // A zero beta_j ends the recurrence: T[0..j+1) is exact and later
  // coefficients would divide by it.  Non-finite coefficients (NaN or Inf in
  // the couplings) stop the problem with error code -2 and no eigenvalues.
  for (int j = p.tested_j; j < n; ++j) {
    if (!std::isfinite(p.host_alpha[j]) || !std::isfinite(p.host_beta[j])) {
      p.final_n = 0;
      p.error_code = -2;
      return;
    }
    if (p.host_beta[j] == 0.0) {
      p.final_n = j + 1;
      return;
    }
  }
  if (p.fixed) {
    if (n >= p.max_iter) p.final_n = p.max_iter;
    return;
  }
  if (n >= 2) {
    double lam = 0.0, lo = 0.0, hi = -1.0;
    if (p.lam_prev == p.lam_prev) {  // bracket from the previous test
      const double drop = std::max(p.lam_drop, 64 * std::numeric_limits<double>::epsilon() * std::fabs(p.lam_prev));
      hi = p.lam_prev + 4 * std::numeric_limits<double>::epsilon() * std::fabs(p.lam_prev);
      lo = p.lam_prev - 4 * drop;
    }
    const bool may = may_be_converged(p.host_alpha, p.host_beta, n, p.fp32, &lam, lo, hi);
    if (p.lam_prev == p.lam_prev) p.lam_drop = std::max(0.0, p.lam_prev - lam);
    p.lam_prev = lam;
    if (may) {
      std::vector<double> a(p.host_alpha, p.host_alpha + n), bb(p.host_beta, p.host_beta + n);
      TValues t = stop_analysis(a, bb, n, p.fp32, p.nev);
      if (nlowest_converged(t, p.nev, p.fp32)) {
        p.final_n = n;
        p.tv = std::move(t);  // conclude() uses this analysis of T[0..n)
        p.tv_n = n;
        return;
      }
    }
  }
  if (n >= p.max_iter) {
    p.final_n = n;
    p.error_code = 1;  // iteration.h:80
  }
}

// After the stopping rule: eigenvalues for FP64 lanes; for FP32 lanes the
// candidates of the first refinement pass.
void conclude(Problem& p) {
  if (p.error_code == -2) {  // non-finite coefficients: nothing to report
    p.ev.clear();
    p.state = kDone;
    return;
  }
  if (p.tv_n != p.final_n) {
    std::vector<double> a(p.host_alpha, p.host_alpha + p.final_n), bb(p.host_beta, p.host_beta + p.final_n);
    p.tv = stop_analysis(a, bb, p.final_n, p.fp32, p.nev);
    p.tv_n = p.final_n;
  }
  if (!p.fp32) {
    p.ev.assign(p.tv.ev.begin(), p.tv.ev.begin() + std::min<size_t>(p.nev, p.tv.ev.size()));
    p.state = kDone;
    return;
  }
  p.state = kRefining;
}

// Ritz coefficients of the next refinement pass (host part).
void prepare_refine(Problem& p) {
  const int m = p.final_n;
  std::vector<double> a(p.host_alpha, p.host_alpha + m), bb(p.host_beta, p.host_beta + m);
  const std::vector<double>& evals = p.tv.all;  // conclude() analysed T[0..m)
  const size_t want = (size_t)std::max(0, p.nev - (int)p.ev.size());
  const size_t cnt = std::min({p.tv.ev.size() - p.next_cand, want, (size_t)p.nvr});
  p.cand.assign(p.tv.index.begin() + p.next_cand, p.tv.index.begin() + p.next_cand + cnt);
  p.next_cand += cnt;
  std::vector<double> z(m);
  for (size_t i = 0; i < cnt; ++i) {
    tridiag_eigenvector(a.data(), bb.data(), m, evals[p.cand[i]], z.data());
    for (int j = 0; j < m; ++j) p.coef_host[i * m + j] = z[j];
  }
}

void ensure_buf(DevBuf& buf, size_t& cap, size_t bytes) {
  if (bytes <= cap) return;
  buf.alloc(std::max<size_t>(bytes, 2 * cap));
  cap = buf.bytes;
}

void ensure_stage(alps_ed_batch b, size_t bytes) {
  if (bytes <= b->stage_cap) return;
  if (b->stage) hip_check(hipHostFree(b->stage), "hipHostFree");
  b->stage = nullptr;
  const size_t cap = std::max<size_t>(bytes, 2 * b->stage_cap);
  hip_check(hipHostMalloc(&b->stage, cap, hipHostMallocDefault), "hipHostMalloc");
  b->stage_cap = cap;
}

// Copies jobs, work items and the job list of a launch group to the device.
struct Group {
  BatchJob* jobs = nullptr;
  BatchItem* hv32 = nullptr;
  BatchItem* hv64 = nullptr;
  BatchItem* up32 = nullptr;
  BatchItem* up64 = nullptr;
  BatchItem* st32 = nullptr;  // start-vector items of the problems new this round
  BatchItem* st64 = nullptr;
  uint32_t* list = nullptr;
  uint32_t* st_list = nullptr;
  uint32_t n_hv32 = 0, n_hv64 = 0, n_up32 = 0, n_up64 = 0, n_list = 0, n_st32 = 0, n_st64 = 0, n_st_list = 0;
  uint32_t seg32 = 0, seg64 = 0;
};

Group upload_group(alps_ed_batch b, const std::vector<Problem*>& run) {
  // Work items, written straight into the pinned stage after one sizing
  // pass (building them in growing vectors and copying cost about 1.7 ms of
  // host time per round for a thousand problems): H.v items (job, segment
  // or block), update items (job, 1024-row chunk), start items of the jobs
  // new this round, then the job list and the new jobs' list.
  Group g;
  const size_t nj = run.size();
  size_t n_hv[2] = {0, 0}, n_up[2] = {0, 0}, n_st[2] = {0, 0}, n_new = 0;
  for (const Problem* p : run) {
    const int f = p->fp32 ? 1 : 0;
    n_hv[f] += p->n_hv;
    n_up[f] += p->n_chunks;
    if (p->j == 0) {
      n_st[f] += p->n_chunks;
      ++n_new;
    }
    if (!p->trans) (p->fp32 ? g.seg32 : g.seg64) = std::max(p->fp32 ? g.seg32 : g.seg64, p->layout.max_seg);
  }
  auto ib = [](size_t n) { return align_up((n + 1) * sizeof(BatchItem)); };
  const size_t bj = align_up(nj * sizeof(BatchJob));
  const size_t bl = align_up((nj + 1) * sizeof(uint32_t));
  const size_t total = bj + ib(n_hv[1]) + ib(n_hv[0]) + ib(n_up[1]) + ib(n_up[0]) + ib(n_st[1]) + ib(n_st[0]) + 2 * bl;
  ensure_buf(b->jobs_d, b->jobs_cap, total);
  ensure_stage(b, total);
  char* s = static_cast<char*>(b->stage);
  char* d = b->jobs_d.as<char>();
  for (size_t i = 0; i < nj; ++i) reinterpret_cast<BatchJob*>(s)[i] = run[i]->job;
  g.jobs = reinterpret_cast<BatchJob*>(d);
  size_t o = bj;
  auto region = [&](size_t n, BatchItem*& dst, uint32_t& count) {
    dst = reinterpret_cast<BatchItem*>(d + o);
    count = (uint32_t)n;
    BatchItem* h = reinterpret_cast<BatchItem*>(s + o);
    o += ib(n);
    return h;
  };
  BatchItem* hv32 = region(n_hv[1], g.hv32, g.n_hv32);
  BatchItem* hv64 = region(n_hv[0], g.hv64, g.n_hv64);
  BatchItem* up32 = region(n_up[1], g.up32, g.n_up32);
  BatchItem* up64 = region(n_up[0], g.up64, g.n_up64);
  BatchItem* st32 = region(n_st[1], g.st32, g.n_st32);
  BatchItem* st64 = region(n_st[0], g.st64, g.n_st64);
  uint32_t* lst = reinterpret_cast<uint32_t*>(s + o);
  g.list = reinterpret_cast<uint32_t*>(d + o);
  g.n_list = (uint32_t)nj;
  o += bl;
  uint32_t* new_list = reinterpret_cast<uint32_t*>(s + o);
  g.st_list = reinterpret_cast<uint32_t*>(d + o);
  g.n_st_list = (uint32_t)n_new;
  uint32_t k_new = 0;
  for (uint32_t i = 0; i < nj; ++i) {
    const Problem& p = *run[i];
    BatchItem*& hv = p.fp32 ? hv32 : hv64;
    BatchItem*& up = p.fp32 ? up32 : up64;
    for (uint32_t sg = 0; sg < p.n_hv; ++sg) *hv++ = {i, sg};
    for (uint32_t c = 0; c < p.n_chunks; ++c) *up++ = {i, c};
    if (p.j == 0) {
      BatchItem*& st = p.fp32 ? st32 : st64;
      for (uint32_t c = 0; c < p.n_chunks; ++c) *st++ = {i, c};
      new_list[k_new++] = i;
    }
    lst[i] = i;
  }
  hip_check(hipMemcpyAsync(d, s, total, hipMemcpyHostToDevice, b->stream), "upload work lists");
  return g;
}

// Per-problem sizes that follow from its layout.
void size_problem(Problem& p, alps_ed_dtype dtype) {
  p.layout = build_layout(p.spec, -1, false, true);  // shared tables stay in L.low / L.flip; lists on the host
  if (p.trans) {
    // the Bloch basis at this momentum; the orbits are shared by the
    // momenta of the same Sz sector
    if (p.layout.D > 0) p.bloch = bloch_sector(*shared_bloch_orbits(p.layout, *p.trans), *p.trans, p.layout.N);
    p.D = p.bloch.rep_state.size();
    p.n = 2 * p.D;
    p.n_hv = (uint32_t)std::max<uint64_t>(1, (p.D + kBlock - 1) / kBlock);
  } else {
    p.D = p.layout.D;
    p.n = p.D;
    p.n_hv = (uint32_t)std::max<size_t>(1, p.layout.seg_hi.size());
  }
  p.fp32 = use_fp32(dtype == ALPS_ED_F32, p.D, p.nev);
  if (p.max_iter <= 0) p.max_iter = (int)std::min<uint64_t>(10 * std::max<uint64_t>(p.D, 1), 1000);
  p.n_chunks = (uint32_t)std::max<uint64_t>(1, (p.n + kBatchChunk - 1) / kBatchChunk);
  p.cap = p.fp32 ? (uint32_t)(p.max_iter + 2) : 3u;
  p.nvr = p.fp32 ? std::min(p.nev, kMaxRitz) : 0;
  p.sized = true;
}

void size_problem(alps_ed_batch b, Problem& p) { size_problem(p, b->cfg.dtype); }

// The device copy of a layout's shared tables, uploaded on first use.
const SharedTables& shared_tables(alps_ed_batch b, const SectorLayout& L) {
  const auto key = std::make_pair(static_cast<const void*>(L.low), static_cast<const void*>(L.flip.get()));
  std::unique_ptr<SharedTables>& slot = b->shared[key];
  if (slot) return *slot;
  slot.reset(new SharedTables);
  SharedTables& t = *slot;
  t.keep = L.flip;
  size_t total = 0;
  auto sz = [&](size_t n, size_t es) {
    const size_t o = total;
    total += align_up(std::max<size_t>(n, 1) * es);
    return o;
  };
  const LowTables& low = *L.low;
  const FlipTables& flip = *L.flip;
  const size_t o1 = sz(low.lo_rank.size(), 4), o2 = sz(low.lo_state.size(), 4), o3 = sz(low.lo_class_off.size(), 4),
               o4 = sz(flip.nbp16.size(), 2), o5 = sz(flip.nbp_off.size(), 4), o6 = sz(flip.xn16.size(), 2),
               o7 = sz(flip.xn_off.size(), 4);
  t.buf.alloc(total);
  char* d = t.buf.as<char>();
  auto up = [&](size_t off, const auto& v) {
    using T = typename std::decay_t<decltype(v)>::value_type;
    if (!v.empty())
      hip_check(hipMemcpyAsync(d + off, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, b->stream),
                "upload shared tables");
    return reinterpret_cast<const T*>(d + off);
  };
  t.lo_rank = up(o1, low.lo_rank);
  t.lo_state = up(o2, low.lo_state);
  t.lo_class_off = up(o3, low.lo_class_off);
  t.nbp16 = up(o4, flip.nbp16);
  t.nbp_off = up(o5, flip.nbp_off);
  t.xn16 = up(o6, flip.xn16);
  t.xn_off = up(o7, flip.xn_off);
  return t;
}

void ensure_tstage(alps_ed_batch b, size_t bytes) {
  if (bytes <= b->tstage_cap) return;
  if (b->tstage) hip_check(hipHostFree(b->tstage), "hipHostFree");
  b->tstage = nullptr;
  const size_t cap = std::max<size_t>(bytes, 2 * b->tstage_cap);
  hip_check(hipHostMalloc(&b->tstage, cap, hipHostMallocDefault), "hipHostMalloc");
  b->tstage_cap = cap;
}

void admit(alps_ed_batch b, std::vector<Problem*>& started) {
  {
    // layouts of the requests at the head of the queue, built in parallel
    std::vector<Problem*> todo;
    for (auto& p : b->pending) {
      if (todo.size() >= 4096) break;
      if (!p->sized) todo.push_back(p.get());
    }
    b->pool->parallel_for(todo.size(), [&](size_t i) { size_problem(b, *todo[i]); });
  }
  std::vector<Problem*> placed;
  while (!b->pending.empty()) {
    if (b->cfg.max_active > 0 && (int)b->active.size() >= b->cfg.max_active) break;
    Problem& p = *b->pending.front();
    if (!p.sized) size_problem(b, p);
    const Footprint f = footprint(p);
    size_t doff = 0, hoff = 0;
    if (f.dev > b->arena.size() || f.host > b->harena.size()) {
      // can never fit: fail this request instead of blocking the queue
      std::unique_ptr<Problem> q = std::move(b->pending.front());
      b->pending.pop_front();
      q->t_admit = now_s();
      q->error_code = -1;
      q->ev.clear();
      finish_problem(b, *q);
      continue;
    }
    if (!b->arena.alloc(f.dev, &doff)) break;
    if (!b->harena.alloc(f.host, &hoff)) {
      b->arena.release(doff, f.dev);
      break;
    }
    std::unique_ptr<Problem> q = std::move(b->pending.front());
    b->pending.pop_front();
    q->dev_off = doff;
    q->dev_bytes = f.dev;
    q->host_off = hoff;
    q->host_bytes = f.host;
    q->t_admit = now_s();
    q->state = kRunning;
    placed.push_back(q.get());
    started.push_back(q.get());
    b->active.push_back(std::move(q));
  }
  if (placed.empty()) return;
  // stage every new problem's tables in pinned memory, with the list of
  // their destinations after them: one upload of the whole stage and one
  // scatter launch (a runtime copy per problem cost more host time than
  // the copies themselves)
  size_t total = 0;
  std::vector<size_t> at;
  for (Problem* p : placed) {
    at.push_back(total);
    total += align_up(problem_table_bytes(*p));
  }
  const size_t items_at = align_up(total);
  const size_t stage_bytes = items_at + placed.size() * sizeof(ScatterItem);
  ensure_tstage(b, stage_bytes);
  char* st = static_cast<char*>(b->tstage);
  std::vector<const SharedTables*> sh(placed.size());
  for (size_t i = 0; i < placed.size(); ++i) sh[i] = &shared_tables(b, placed[i]->layout);
  b->pool->parallel_for(placed.size(), [&](size_t i) { place(b, *placed[i], st + at[i], *sh[i]); });
  ScatterItem* items = reinterpret_cast<ScatterItem*>(st + items_at);
  for (size_t i = 0; i < placed.size(); ++i)
    items[i] = ScatterItem{at[i], problem_table_bytes(*placed[i]), static_cast<char*>(b->dev_base) + placed[i]->dev_off};
  ensure_buf(b->tdev, b->tdev_cap, stage_bytes);
  hip_check(hipMemcpyAsync(b->tdev.p, st, stage_bytes, hipMemcpyHostToDevice, b->stream), "upload tables");
  hip_check(launch_scatter_copy(b->tdev.as<unsigned char>(), reinterpret_cast<const ScatterItem*>(b->tdev.as<char>() + items_at),
                                (uint32_t)placed.size(), b->stream),
            "scatter tables");
}

void retire(alps_ed_batch b) {
  std::vector<std::unique_ptr<Problem>> keep;
  for (auto& p : b->active) {
    if (p->state == kDone) {
      finish_problem(b, *p);
      b->arena.release(p->dev_off, p->dev_bytes);
      b->harena.release(p->host_off, p->host_bytes);
    } else {
      keep.push_back(std::move(p));
    }
  }
  b->active.swap(keep);
}

// Refinement pass for every problem in kRefining: Ritz vectors from the
// kept basis, H y and the Rayleigh quotients (all FP64 accumulation).
void refine(alps_ed_batch b) {
  std::vector<Problem*> ref;
  for (auto& p : b->active)
    if (p->state == kRefining) ref.push_back(p.get());
  if (ref.empty()) return;
  b->pool->parallel_for(ref.size(), [&](size_t i) { prepare_refine(*ref[i]); });
  // drop problems with no candidate left
  std::vector<Problem*> work;
  for (Problem* p : ref) {
    if (p->cand.empty()) {
      p->state = kDone;
      continue;
    }
    work.push_back(p);
  }
  if (work.empty()) return;
  // Ritz jobs and, per candidate rank v, the plain H.v jobs for H y
  std::vector<RitzJob> rj(work.size());
  std::vector<BatchItem> ri;
  for (uint32_t i = 0; i < work.size(); ++i) {
    Problem& p = *work[i];
    RitzJob& r = rj[i];
    r.vec = p.job.vec;
    r.stride = p.job.stride;
    r.D = (uint32_t)p.n;
    r.n_chunks = p.n_chunks;
    r.m = p.final_n;
    r.nv = (int)p.cand.size();
    r.scale = p.job.scale;
    r.coef = p.coef;
    r.out = p.ritz;
    r.part = mapped_dev(b, p.ritz_part);
    for (uint32_t c = 0; c < p.n_chunks; ++c) ri.push_back({i, c});
  }
  size_t max_nv = 0;
  for (Problem* p : work) max_nv = std::max(max_nv, p->cand.size());
  std::vector<std::vector<PlainHvJob>> hj(max_nv);
  std::vector<std::vector<BatchItem>> hi(max_nv);
  std::vector<std::vector<size_t>> who(max_nv);  // positions in work
  std::vector<uint32_t> max_seg(max_nv, 0);
  for (size_t v = 0; v < max_nv; ++v)
    for (size_t wi = 0; wi < work.size(); ++wi) {
      Problem* p = work[wi];
      if (v >= p->cand.size()) continue;
      PlainHvJob j;
      j.L = p->job.L;
      j.Bl = p->job.Bl;
      j.x = p->ritz + v * p->n;
      j.y = p->hy;
      j.part = mapped_dev(b, p->hv_part);
      const uint32_t ji = (uint32_t)hj[v].size();
      hj[v].push_back(j);
      who[v].push_back(wi);
      for (uint32_t sg = 0; sg < p->n_hv; ++sg) hi[v].push_back({ji, sg});
      if (!p->trans) max_seg[v] = std::max(max_seg[v], p->layout.max_seg);
    }
  // One pinned stage holds every upload of the pass: the Ritz coefficients
  // (scattered to each problem's coefficient buffer), the Ritz jobs and
  // items, and the H.v jobs and items of every rank; one copy brings it to
  // the device (a runtime copy, allocation and free per problem and per rank
  // cost more host time than the work).
  size_t o = 0;
  auto reserve = [&](size_t bytes) {
    const size_t at = o;
    o += align_up(bytes);
    return at;
  };
  std::vector<size_t> coef_at(work.size());
  for (size_t i = 0; i < work.size(); ++i) coef_at[i] = reserve(work[i]->cand.size() * work[i]->final_n * sizeof(double));
  const size_t items_at = reserve(work.size() * sizeof(ScatterItem));
  const size_t rj_at = reserve(rj.size() * sizeof(RitzJob));
  const size_t ri_at = reserve(ri.size() * sizeof(BatchItem));
  std::vector<size_t> hj_at(max_nv), hi_at(max_nv);
  for (size_t v = 0; v < max_nv; ++v) {
    hj_at[v] = reserve(hj[v].size() * sizeof(PlainHvJob));
    hi_at[v] = reserve(hi[v].size() * sizeof(BatchItem));
  }
  if (o > b->rstage_cap) {
    if (b->rstage) hip_check(hipHostFree(b->rstage), "hipHostFree");
    b->rstage = nullptr;
    b->rstage_cap = std::max(o, 2 * b->rstage_cap);
    hip_check(hipHostMalloc(&b->rstage, b->rstage_cap, hipHostMallocDefault), "hipHostMalloc");
  }
  ensure_buf(b->rdev, b->rdev_cap, o);
  char* st = static_cast<char*>(b->rstage);
  char* dv = b->rdev.as<char>();
  ScatterItem* items = reinterpret_cast<ScatterItem*>(st + items_at);
  for (size_t i = 0; i < work.size(); ++i) {
    const size_t bytes = work[i]->cand.size() * work[i]->final_n * sizeof(double);
    std::memcpy(st + coef_at[i], work[i]->coef_host, bytes);
    items[i] = ScatterItem{coef_at[i], bytes, work[i]->coef};
  }
  std::memcpy(st + rj_at, rj.data(), rj.size() * sizeof(RitzJob));
  std::memcpy(st + ri_at, ri.data(), ri.size() * sizeof(BatchItem));
  for (size_t v = 0; v < max_nv; ++v) {
    std::memcpy(st + hj_at[v], hj[v].data(), hj[v].size() * sizeof(PlainHvJob));
    std::memcpy(st + hi_at[v], hi[v].data(), hi[v].size() * sizeof(BatchItem));
  }
  hip_check(hipMemcpyAsync(dv, st, o, hipMemcpyHostToDevice, b->stream), "upload refinement jobs");
  hip_check(launch_scatter_copy(reinterpret_cast<const unsigned char*>(dv), reinterpret_cast<const ScatterItem*>(dv + items_at),
                                (uint32_t)work.size(), b->stream),
            "scatter coefficients");
  hip_check(launch_batch_ritz<float>(reinterpret_cast<RitzJob*>(dv + rj_at), reinterpret_cast<BatchItem*>(dv + ri_at),
                                     (uint32_t)ri.size(), b->stream),
            "batch ritz");
  // H y for each candidate, one plain H.v group per candidate rank (each
  // problem's y and partials are reused by the next rank)
  std::vector<std::vector<double>> yhy(work.size());
  for (size_t v = 0; v < max_nv; ++v) {
    hip_check(launch_batch_plain_hv(reinterpret_cast<PlainHvJob*>(dv + hj_at[v]), reinterpret_cast<BatchItem*>(dv + hi_at[v]),
                                    (uint32_t)hi[v].size(), max_seg[v], b->stream),
              "plain hv");
    hip_check(hipStreamSynchronize(b->stream), "sync");
    for (size_t wi : who[v]) {
      const Problem& p = *work[wi];
      double s = 0.0;
      for (size_t i = 0; i < p.n_hv; ++i) s += p.hv_part[i];
      yhy[wi].push_back(s);
    }
  }
  for (size_t i = 0; i < work.size(); ++i) {
    Problem& p = *work[i];
    for (size_t v = 0; v < p.cand.size(); ++v) {
      double yy = 0.0;
      for (uint32_t c = 0; c < p.n_chunks; ++c) yy += p.ritz_part[v * p.n_chunks + c];
      const double rq = yhy[i][v] / yy;
      bool dup = false;
      for (double e : p.ev)
        if (std::fabs(rq - e) <= 1e-9 * std::max(1.0, std::fabs(e))) dup = true;
      if (!dup) p.ev.push_back(rq);
    }
    std::sort(p.ev.begin(), p.ev.end());
    p.cand.clear();
    if ((int)p.ev.size() < p.nev && p.next_cand >= p.tv.ev.size() && !p.tv.complete) {
      // more candidates than the stopping test listed: the full analysis
      // (whose first entries are the same) lists them all
      std::vector<double> a(p.host_alpha, p.host_alpha + p.final_n), bb(p.host_beta, p.host_beta + p.final_n);
      p.tv = analyze_prefix(a, bb, p.final_n, p.fp32);
    }
    if ((int)p.ev.size() >= p.nev || p.next_cand >= p.tv.ev.size()) p.state = kDone;
  }
}

// D = 0 and D = 1 sectors need no Lanczos (sparsediag.h:101-102,120-124).
// A momentum sector's host basis is on the device by now and is released.
void settle_trivial(const std::vector<Problem*>& started) {
  for (Problem* p : started) {
    if (p->D == 0) {
      p->state = kDone;
    } else if (p->D == 1) {
      p->ev.assign(1, p->trans ? bloch_single_element(p->layout, p->bloch) : single_state_energy(p->spec));
      p->final_n = 0;
      p->state = kDone;
    }
    if (p->trans) p->bloch = BlochSector();
  }
}

// One round.  While the device runs this round's steps the host admits new
// requests (they start next round) and tests the stopping rule on the
// previous round's alpha/beta, so neither costs device time; a problem found
// converged has then run one more round of steps, which changes nothing it
// reports.  When nothing runs, requests are admitted first and start at once.
void round(alps_ed_batch b) {
  ++b->rounds;
  const double t0 = now_s();
  auto running = [&](std::vector<Problem*>& run, std::vector<Problem*>& test) {
    run.clear();
    test.clear();
    for (auto& p : b->active)
      if (p->state == kRunning) {
        run.push_back(p.get());
        if (p->j > p->tested_j) test.push_back(p.get());
      }
  };
  std::vector<Problem*> run, test, started;
  running(run, test);
  bool admitted = false;
  if (run.empty()) {
    admit(b, started);
    settle_trivial(started);
    running(run, test);
    admitted = true;
  }
  const double t1 = now_s();
  b->t_admit += t1 - t0;
  const int S = std::max(1, b->cfg.steps_per_round);
  if (!run.empty()) {
    for (Problem* p : run) p->job.j0 = p->j;
    Group g = upload_group(b, run);
    if (g.n_st_list) {
      hip_check(launch_batch_start<float>(g.jobs, g.st32, g.n_st32, b->stream), "batch start");
      hip_check(launch_batch_start<double>(g.jobs, g.st64, g.n_st64, b->stream), "batch start");
      hip_check(launch_batch_finish(g.jobs, g.st_list, g.n_st_list, 0, kBatchStart, b->stream), "batch finish");
    }
    auto steps = [&] {
      for (int t = 0; t < S; ++t) {
        hip_check(launch_batch_hv<float>(g.jobs, g.hv32, g.n_hv32, t, g.seg32, b->stream), "batch hv");
        hip_check(launch_batch_hv<double>(g.jobs, g.hv64, g.n_hv64, t, g.seg64, b->stream), "batch hv");
        hip_check(launch_batch_finish(g.jobs, g.list, g.n_list, t, kBatchAlpha, b->stream), "batch finish");
        hip_check(launch_batch_update<float>(g.jobs, g.up32, g.n_up32, t, b->stream), "batch update");
        hip_check(launch_batch_update<double>(g.jobs, g.up64, g.n_up64, t, b->stream), "batch update");
        hip_check(launch_batch_finish(g.jobs, g.list, g.n_list, t, kBatchBeta, b->stream), "batch finish");
      }
    };
    // The launches of a round are fixed by the work-list addresses and
    // counts and S (each job's step base j0 is read from device memory), so
    // a round that repeats the previous one replays them from a graph.
    bool replayed = false;
    if (b->cfg.graphs && g.n_st_list == 0) {
      const std::vector<uint64_t> key{(uint64_t)g.jobs, (uint64_t)g.hv32, (uint64_t)g.hv64, (uint64_t)g.up32,
                                      (uint64_t)g.up64, (uint64_t)g.list, g.n_hv32, g.n_hv64, g.n_up32, g.n_up64,
                                      g.n_list, g.seg32, g.seg64, (uint64_t)S};
      if (key == b->last_key) {
        if (!b->graph_exec || key != b->graph_key) {
          if (b->graph_exec) hip_check(hipGraphExecDestroy(b->graph_exec), "graph destroy");
          b->graph_exec = nullptr;
          hipGraph_t graph = nullptr;
          hip_check(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal), "graph capture");
          steps();
          hip_check(hipStreamEndCapture(b->stream, &graph), "graph capture");
          const hipError_t e = hipGraphInstantiate(&b->graph_exec, graph, nullptr, nullptr, 0);
          (void)hipGraphDestroy(graph);
          hip_check(e, "graph instantiate");
          b->graph_key = key;
          ++b->graph_captures;
        }
        hip_check(hipGraphLaunch(b->graph_exec, b->stream), "graph launch");
        ++b->graph_rounds;
        replayed = true;
      }
      b->last_key = key;
    } else {
      b->last_key.clear();
    }
    if (!replayed) steps();
  }
  // host work overlapping the steps: admission (table copies queue behind
  // the steps on the stream), then the stopping tests on results of earlier
  // rounds (entries below p->j are final)
  const double t2 = now_s();
  if (!admitted) {
    admit(b, started);
    settle_trivial(started);
  }
  const double t3 = now_s();
  b->t_admit += t3 - t2;
  b->pool->parallel_for(test.size(), [&](size_t i) {
    Problem& p = *test[i];
    decide(p);
    p.tested_j = p.j;
    if (p.final_n >= 0) conclude(p);
  });
  const double t4 = now_s();
  b->t_decide += t4 - t3;
  if (!run.empty()) {
    hip_check(hipStreamSynchronize(b->stream), "sync");
    b->steps += S;
    for (Problem* p : run) {
      const int j1 = std::min(p->max_iter, p->j + S);
      b->hv_rows[p->fp32 ? 1 : 0] += p->D * (uint64_t)(j1 - p->j);
      p->j = j1;
    }
  }
  const double t5 = now_s();
  b->t_steps += t5 - t4;
  refine(b);
  b->t_refine += now_s() - t5;
  retire(b);
  hip_check(hipStreamSynchronize(b->stream), "sync");  // staging buffers are reused next round
}

// Translations of a momentum request, every one a permutation of the sites.
std::shared_ptr<const Translations> checked_translations(const alps_ed_model* model, int n_t, const int* perm,
                                                         const double* re, const double* im) {
  require(model->n_sites >= 1 && model->n_sites <= 32, "momentum sectors support 1 to 32 sites");
  auto t = std::make_shared<Translations>(read_translations(model->n_sites, n_t, perm, re, im));
  for (const auto& q : t->perm) {
    std::vector<int> seen(model->n_sites, 0);
    for (int s : q) {
      require(s >= 0 && s < model->n_sites && !seen[s], "translation is not a permutation of the sites");
      seen[s] = 1;
    }
  }
  return t;
}

uint64_t submit_one(alps_ed_batch b, const alps_ed_model* m, const alps_ed_solve_params* prm,
                    std::shared_ptr<const Translations> trans = nullptr) {
  require(prm != nullptr, "params is null");
  std::unique_ptr<Problem> p(new Problem);
  p->spec = model_spec(m);
  p->trans = std::move(trans);
  if (p->spec.n_up >= 0 && p->spec.n_up <= p->spec.n_sites)
    require(binom(p->spec.n_sites, p->spec.n_up) < 0xffffffffULL, "sector dimension exceeds 2^32-1");
  p->max_iter = prm->max_iter > 0 ? prm->max_iter : 0;  // default set when sized
  p->nev = prm->nev > 0 ? prm->nev : 1;
  require(p->nev <= 64, "nev must be <= 64");
  p->fixed = prm->fixed_iterations != 0;
  p->seed = prm->seed;
  p->ticket = b->next_ticket++;
  p->t_submit = now_s();
  const uint64_t t = p->ticket;
  b->pending.push_back(std::move(p));
  return t;
}

}  // namespace

extern "C" {

int alps_ed_batch_create(const alps_ed_batch_config* cfg, alps_ed_batch* out) {
  return guarded([&] {
    require(cfg && out, "null argument");
    require(cfg->dtype == ALPS_ED_F32 || cfg->dtype == ALPS_ED_F64, "unknown dtype");
    *out = nullptr;
    std::unique_ptr<alps_ed_batch_s> b(new alps_ed_batch_s);
    b->cfg = *cfg;
    if (b->cfg.steps_per_round <= 0) b->cfg.steps_per_round = 8;
    if (b->cfg.host_threads <= 0) b->cfg.host_threads = 8;
    size_t budget = cfg->hbm_bytes;
    if (budget == 0) {
      size_t fr = 0, tot = 0;
      hip_check(hipMemGetInfo(&fr, &tot), "hipMemGetInfo");
      budget = fr / 2;
    }
    budget = align_up(budget);
    hip_check(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "hipStreamCreate");
    hip_check(hipMalloc(&b->dev_base, budget), "hipMalloc arena");
    b->arena.init(budget);
    const size_t hbytes = align_up(cfg->host_bytes ? cfg->host_bytes : ((size_t)256 << 20));
    void* hp = nullptr;
    hip_check(hipHostMalloc(&hp, hbytes, hipHostMallocMapped), "hipHostMalloc");
    b->host_base = static_cast<char*>(hp);
    void* dp = nullptr;
    hip_check(hipHostGetDevicePointer(&dp, hp, 0), "hipHostGetDevicePointer");
    b->host_dev = static_cast<char*>(dp);
    b->harena.init(hbytes);
    b->pool.reset(new Pool(b->cfg.host_threads - 1));
    *out = b.release();
  });
}

int alps_ed_batch_destroy(alps_ed_batch b) {
  return guarded([&] { delete b; });
}

int alps_ed_batch_submit(alps_ed_batch b, const alps_ed_model* model, const alps_ed_solve_params* params,
                         uint64_t* ticket) {
  return guarded([&] {
    require(b != nullptr, "batch is null");
    const uint64_t t = submit_one(b, model, params);
    if (ticket) *ticket = t;
  });
}

int alps_ed_batch_submit_momentum(alps_ed_batch b, const alps_ed_model* model, int n_translations, const int* perm,
                                  const double* phase_re, const double* phase_im, const alps_ed_solve_params* params,
                                  uint64_t* ticket) {
  return guarded([&] {
    require(b != nullptr && model != nullptr && ticket != nullptr, "null argument");
    *ticket = submit_one(b, model, params, checked_translations(model, n_translations, perm, phase_re, phase_im));
  });
}

int alps_ed_batch_progress(alps_ed_batch b, int* n_active, int* n_pending) {
  return guarded([&] {
    require(b != nullptr, "batch is null");
    if (!b->active.empty() || !b->pending.empty()) round(b);
    if (n_active) *n_active = (int)b->active.size();
    if (n_pending) *n_pending = (int)b->pending.size();
  });
}

int alps_ed_batch_run(alps_ed_batch b) {
  return guarded([&] {
    require(b != nullptr, "batch is null");
    while (!b->active.empty() || !b->pending.empty()) round(b);
  });
}

int alps_ed_batch_completed(alps_ed_batch b, uint64_t* tickets, int max, int* n) {
  return guarded([&] {
    require(b != nullptr && n != nullptr, "null argument");
    const int cnt = (int)std::min<size_t>(b->completed.size(), (size_t)std::max(0, max));
    for (int i = 0; i < cnt; ++i) tickets[i] = b->completed[i];
    b->completed.erase(b->completed.begin(), b->completed.begin() + cnt);
    *n = cnt;
  });
}

int alps_ed_batch_result(alps_ed_batch b, uint64_t ticket, alps_ed_solve_result* result, int* ready) {
  return guarded([&] {
    require(b != nullptr && result != nullptr, "null argument");
    auto it = b->results.find(ticket);
    if (ready) *ready = it != b->results.end();
    if (it == b->results.end()) {
      std::memset(result, 0, sizeof(*result));
      return;
    }
    *result = it->second;
  });
}

int alps_ed_batch_stats(alps_ed_batch b, alps_ed_batch_status* s) {
  return guarded([&] {
    require(b != nullptr && s != nullptr, "null argument");
    s->n_active = (int)b->active.size();
    s->n_pending = (int)b->pending.size();
    s->hbm_used = b->arena.used();
    s->hbm_bytes = b->arena.size();
    s->rounds = b->rounds;
    s->steps = b->steps;
    s->solves_done = b->solves_done;
    s->seconds_admit = b->t_admit;
    s->seconds_steps = b->t_steps;
    s->seconds_decide = b->t_decide;
    s->seconds_refine = b->t_refine;
    s->hv_rows_f32 = b->hv_rows[1];
    s->hv_rows_f64 = b->hv_rows[0];
    s->graph_rounds = b->graph_rounds;
    s->graph_captures = b->graph_captures;
  });
}

int alps_ed_batch_footprint(const alps_ed_model* model, const alps_ed_solve_params* params, alps_ed_dtype dtype,
                            uint64_t* bytes) {
  return guarded([&] {
    require(params && bytes, "null argument");
    Problem p;
    p.spec = model_spec(model);
    p.nev = params->nev > 0 ? params->nev : 1;
    p.max_iter = params->max_iter > 0 ? params->max_iter : 0;
    size_problem(p, dtype);
    *bytes = footprint(p).dev;
  });
}

int alps_ed_batch_footprint_momentum(const alps_ed_model* model, int n_translations, const int* perm,
                                     const double* phase_re, const double* phase_im,
                                     const alps_ed_solve_params* params, alps_ed_dtype dtype, uint64_t* bytes) {
  return guarded([&] {
    require(model && params && bytes, "null argument");
    Problem p;
    p.spec = model_spec(model);
    p.trans = checked_translations(model, n_translations, perm, phase_re, phase_im);
    p.nev = params->nev > 0 ? params->nev : 1;
    p.max_iter = params->max_iter > 0 ? params->max_iter : 0;
    size_problem(p, dtype);
    *bytes = footprint(p).dev;
  });
}

}  // extern "C"
