// ed_kernels.hpp -- device data layout and kernel launchers (host interface).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "ed_layout.hpp"

namespace alps_ed {

// Device-resident sector layout; all pointers are device memory.
struct DevLayout {
  int N, k, B, NH;
  uint32_t D;
  uint32_t n_seg, max_seg;
  int n_ll_p, n_hh_p, n_xb_p, n_bonds;  // padded flip-bond counts (table strides)
  int n_ll, n_hh, n_xb;                 // flip bonds per class
  const uint32_t* lo_rank;
  const uint32_t* lo_state;
  const uint32_t* lo_class_off;
  const uint32_t* hi_base;
  const uint32_t* seg_hi;     // ascending (unrank)
  const uint32_t* seg_start;
  const uint32_t* proc_hi;    // H.v processing order
  const double* diag_lo_cls;  // low-word diagonal in lo_state order
  const double* diag_hi;
  const uint16_t* nbp16;
  const uint32_t* nbp_off;
  const uint16_t* xn16;
  const uint32_t* xn_off;
  const double* ll_t;
  const uint32_t* hh_off;     // [2^NH + 1] per high word: its flipping high-high bonds
  const HHEntry* hh_ent;      //   (neighbour segment, coefficient), padded with 0-coefficient entries
  const XBondConst* xb_c;     // [n_xb_p] cross-bond constants
  const uint32_t* xb_nb;      // [2^NH * n_xb_p] first row of the cross-bond neighbour segment
  const SegRec* seg_rec;      // [n_seg] per processing index (single operators; the batch looks them up)
  // reference-order description for the element enumeration
  const double* site_dn;
  const double* site_up;
  const uint32_t* bond_p1;
  const uint32_t* bond_p2;
  const double* bond_par;
  const double* bond_anti;
  const double* bond_t;
};

// A momentum sector (Bloch basis, ed_bloch.hpp) on the device.  Vectors of
// the sector are complex, stored as 2 * dk reals with the real and imaginary
// parts of each element side by side (Re v_0, Im v_0, Re v_1, ...: the
// layout of an array of std::complex).
struct DevBloch {
  uint32_t dk;                // kept representatives
  const uint32_t* rep_state;  // [dk]
  const double* norm;         // [dk]
  const uint64_t* full_pk;    // [D] per rank of the fixed-Sz sector: packed
                              // representative, translation, normalization index
  const double* norm_tab;     // distinct normalizations
  const double* ph_re;        // [translations]
  const double* ph_im;
  // The flips of each kept representative, listed once on the host
  // (bloch_flip_rows, ed_bloch.hpp); null where not built (the batch, which
  // uses bloch_row).  Entry = target representative | coefficient code <<
  // sh_c; the code indexes the complex coefficient t_b * norm_a * phase_t
  // (coef[2 code], coef[2 code + 1]), which the row scales by 1 / norm_i.
  const uint32_t* row_ptr;    // [dk + 1]
  const uint32_t* ent;
  const double* diag;         // [dk] the diagonal element of each row
  const double* inv_norm;     // [dk] 1 / norm
  const double* coef;         // per coefficient code: re, im
  uint32_t sh_c, mask_a;
};

// Reductions without atomics: the producing kernel stores one partial sum
// per block; a consuming kernel adds the partials in index order (every
// block of the consumer computes the same value, so results are
// deterministic and need no grid synchronization).
struct Partials {
  const double* p;  // nullptr when absent
  uint32_t n;
};

constexpr int kBlock = 256;
// H.v kernels on FP32 vectors are compiled for at least 8 waves per SIMD
// (at most 64 registers): more waves in flight hide the gather latency.
// The FP64 instances keep the default (8 would spill).
template <class V>
struct HvWaves {
  static constexpr int value = sizeof(V) == 4 ? 8 : 1;
};
// FP32 storage of the Lanczos vectors is used, when asked for, only where it
// pays: sectors of at least kFp32MinDim states (in small sectors the Krylov
// space is a large part of the sector).  Without reorthogonalization (the
// batch) also at most kFp32MaxNev requested eigenvalues: for more, the copies
// of converged levels that FP32 rounding creates early delay the higher
// levels (about twice the FP64 iteration count for four levels), which costs
// more than the halved bytes save.  Single solves reorthogonalize FP32
// recurrences of more than kFp32MaxNev levels partially (ReorthDev), which
// keeps the copies away, so they store any number of levels in FP32 (for
// one or two levels the copies do not delay the stopping test and the
// reorthogonalization would only cost time).  Elsewhere the vectors are FP64.
constexpr uint64_t kFp32MinDim = 1000;
constexpr int kFp32MaxNev = 2;
inline bool use_fp32(bool asked, uint64_t dim, int nev) { return asked && dim >= kFp32MinDim && nev <= kFp32MaxNev; }
inline bool use_fp32_reorth(bool asked, uint64_t dim) { return asked && dim >= kFp32MinDim; }
constexpr int kMaxRitz = 8;     // Ritz vectors per pass of launch_ritz
// H.v rows per thread: a segment of up to kRows*kBlock rows is one workgroup
constexpr int kRows = 4;

// Number of blocks of the streaming kernels for a vector of length n.
inline unsigned stream_blocks(uint64_t n) {
  uint64_t b = (n + (uint64_t)kBlock * 4 - 1) / ((uint64_t)kBlock * 4);
  if (b < 1) b = 1;
  if (b > 1024) b = 1024;
  return (unsigned)b;
}

// Blocks of the fused update: as the streaming kernels (at most 1024).  Every
// block also re-reads all partials of the step; on the L=24 ring 1024 blocks
// were 4 % faster per Lanczos step than 256 blocks, and than a separate
// one-block scalar kernel before a full-width update.
inline unsigned fused_blocks(uint64_t n) { return stream_blocks(n); }

// Blocks of the H.v kernel: one per segment.
inline unsigned hv_blocks(const DevLayout& L) { return L.n_seg ? L.n_seg : 1; }

// H.v schedules for FP32 vectors: the waves per SIMD the kernel is compiled
// for and the chunk sizes of its low-low, high-high and cross flip loops.
// Schedule 0 is the default (FP64 vectors always use it); an operator of a
// large FP32 sector picks its schedule by timing them all (the autotuner in
// ed_capi.hip).  Every schedule adds the bonds in the same order and forms
// the dot partials the same way, so all give identical results.
struct HvSchedule {
  int waves, cll, chh, cx;
};
constexpr int kHvSchedules = 6;
constexpr HvSchedule kHvSchedule[kHvSchedules] = {{8, 4, 4, 2}, {8, 4, 8, 2}, {6, 4, 8, 2},
                                                  {6, 4, 4, 2}, {8, 8, 4, 2}, {6, 8, 8, 4}};

// y = H x; with dot_out set, per-block partials of sum_i x_i (H x)_i (FP64,
// before y is rounded to the storage type).
template <class V>
hipError_t launch_hv(const DevLayout& L, const V* x, V* y, double* dot_out, hipStream_t st, int schedule = 0);

// Lanczos scalars between the kernels of one step (one block):
//   mode kAlpha: alpha = scale^2 * sum(in)            -> out[0]
//   mode kBeta:  beta = sqrt(sum(in)), scale = 1/beta -> out[0], out2[0]
enum { kSum = 0, kNorm = 1, kAlpha = 2 };
// out_h, when set (host-mapped memory), receives a copy of out[0].
hipError_t launch_finish(Partials in, const double* scale, double* out, double* out2, int mode, hipStream_t st,
                         double* out_h = nullptr);

// Lanczos update: r = s_cur*w - alpha*s_cur*u_cur - beta_prev*s_prev*u_prev
// stored into `out` (may alias w or u_prev); per-block partials of |r|^2
// into norm_out.  Scalars are read from device memory; beta_prev == nullptr
// drops the last term; null s_cur / s_prev mean 1.
template <class V>
hipError_t launch_update(uint64_t n, const V* w, const V* u_cur, const V* u_prev, V* out, const double* alpha,
                         const double* s_cur, const double* beta_prev, const double* s_prev, double* norm_out,
                         hipStream_t st);

// Partial reorthogonalization (Simon, "The Lanczos algorithm with partial
// reorthogonalization", Math. Comp. 42 (1984) 115-142) of a recurrence whose
// basis u_0..u_j is kept in HBM (v_k = scale_k u_k).  After the update of
// step j has stored r = beta_j v_{j+1} as u_{j+1}, the first Gram-Schmidt
// kernel extends the estimates omega_{j+1,k} ~ v_{j+1}.v_k by the omega
// recurrence in FP64 (every block the same way; block 0 stores them) and
// decides: reorthogonalize when max_k |omega_{j+1,k}| > sqrt(eps) (eps =
// unit roundoff of the vector storage), and again at the next step (the pair
// rule).  The Gram-Schmidt kernels do their work only when the step is
// flagged (uniform early exit otherwise): per-block partials of u_k.r, the
// coefficients gamma_k = scale_k^2 u_k.r, and r -= sum_k gamma_k u_k with new
// norm partials of r in place of the update's.  Every sum is in a fixed
// order, so the recurrence stays deterministic.
struct ReorthDev {
  double* omega;            // three rows of `stride`: row of step j at (j % 3) * stride
  uint32_t stride;          // max_steps + 2
  int* flag;                // per step: 0 none, 1 threshold crossed, 2 second of a pair
  double* gamma;            // per kept vector: the Gram-Schmidt coefficient
  double* dot_part;         // kReorthDotBlocks partials per kept vector
  double* hnorm;            // per step: running estimate of |T|
  const double* alpha;      // the recurrence's alpha and beta
  const double* beta;
  const void* const* slots; // device addresses of u_0, u_1, ...
  double eps;               // unit roundoff of the vector storage
};
constexpr int kReorthDotBlocks = 1024;
// The omega step and Gram-Schmidt pass of step j on r = u_{j+1} (three
// launches; norm_out holds the update's fused_blocks(n) norm partials of r
// and receives those of the reorthogonalized r when the step is flagged).
template <class V>
hipError_t launch_reorth(uint64_t n, int j, V* r, const double* scale, const ReorthDev& R, double* norm_out,
                         hipStream_t st);

// The per-high-word lists of a layout on the device: launch_hh_count writes
// off[0..nhi] (the prefix sums of each high word's flipping high-high bond
// count, rounded up to kHHList; 0xffffffff once they reach 2^32 - 1) and
// xb_nb[nhi * n_xb_p]; launch_hh_fill then writes the off[nhi] entries.  The
// same lists build_layout makes on the host with host_lists.
hipError_t launch_hh_count(uint32_t nhi, const uint32_t* hi_base, const uint32_t* mask, int nb, uint32_t* off,
                           const uint32_t* hbit, int n_xb_p, uint32_t* xb_nb, hipStream_t st);
hipError_t launch_hh_fill(uint32_t nhi, const uint32_t* hi_base, const uint32_t* mask, const double* t, int nb,
                          const uint32_t* off, HHEntry* ent, hipStream_t st);

// seg_rec[n_seg] (SegRec, one per processing index) of a layout whose tables
// and lists are on the device.
hipError_t launch_seg_rec(const DevLayout& L, SegRec* out, hipStream_t st);

// The update with its scalars (one launch instead of three): scale_j =
// 1/sqrt(sum(norm_prev)), beta_{j-1} = sqrt(sum(norm_prev)) (when beta_jm1 is
// set; it is null at j = 0), alpha_j = scale_j^2 * sum(hv_part); then
// r = scale_j w - alpha_j scale_j u_cur - beta_{j-1} scale_{j-1} u_prev into
// `out` with per-block partials of |r|^2 into norm_out (fused_blocks(n) of
// them).  The scalars are bit-identical to launch_finish + launch_update.
// alpha_h / beta_h, when set (host-mapped memory), receive copies of alpha_j
// and beta_{j-1}.
template <class V>
hipError_t launch_update_fused(uint64_t n, const V* w, const V* u_cur, const V* u_prev, V* out, Partials hv_part,
                               Partials norm_prev, double* alpha_j, double* beta_jm1, double* scale_j,
                               const double* scale_jm1, double* norm_out, hipStream_t st, double* alpha_h = nullptr,
                               double* beta_h = nullptr);

// per-block partials of sum x_i y_i
template <class V>
hipError_t launch_dot(uint64_t n, const V* x, const V* y, double* partials, hipStream_t st);
// v *= 1/sqrt(sum(norm_in)); block 0 stores sqrt(sum) to b_out
template <class V>
hipError_t launch_normalize(uint64_t n, V* v, Partials norm_in, double* b_out, hipStream_t st);

template <class V>
hipError_t launch_fill_random(uint64_t n, V* v, uint64_t seed, double lo, double hi, hipStream_t st);

// Copies of many blocks out of one staged device buffer: item i copies
// items[i].bytes from src + items[i].src to items[i].dst (one launch
// instead of a runtime copy per block).
struct ScatterItem {
  uint64_t src, bytes;
  void* dst;
};
hipError_t launch_scatter_copy(const unsigned char* src, const ScatterItem* items, uint32_t n, hipStream_t st);
template <class V>
hipError_t launch_fill(uint64_t n, V* v, double value, hipStream_t st);
// y = a*x + b*y
template <class V>
hipError_t launch_axpby(uint64_t n, V* y, double a, const V* x, double b, hipStream_t st);
template <class VD, class VS>
hipError_t launch_convert(uint64_t n, VD* dst, const VS* src, const double* scale, hipStream_t st);
// out_i = sum_{j<m} coef[i*m+j] * scale[j] * slots[j], i < nvec <= kMaxRitz
template <class V>
hipError_t launch_ritz(uint64_t n, const V* const* slots, const double* scale, int m, int nvec,
                       const double* coef, double* const* outs, hipStream_t st);

// Diagonal spin moments of v (ed_measure.hip): out[0] = sum v_r^2,
// out[1 + i] = sum v_r^2 Sz_i(r), out[1 + N + i*N + j] = sum v_r^2 Sz_i Sz_j;
// `stats` holds n_seg * moment_stat_count(L) doubles of scratch.
int moment_stat_count(const DevLayout& L);
template <class V>
hipError_t launch_spin_moments(const DevLayout& L, const V* v, double* stats, double* out, hipStream_t st);
// Flip amplitudes out[p] = <v| S+_i S-_j |v> for masks[2p] = bit of site j,
// masks[2p+1] = bit of site i (i != j); `part` holds
// n_pairs * flip_pair_blocks(L) doubles of scratch.
uint32_t flip_pair_blocks(const DevLayout& L);
template <class V>
hipError_t launch_flip_pairs(const DevLayout& L, const V* v, const uint32_t* masks, int n_pairs, double* part,
                             double* out, hipStream_t st);

// Lanes per row of the listed-flips Bloch kernel (4 measured fastest on the
// L=24 ring's k = 0 sector: 6.05 us against 6.19 us with 2 and 7.52 us with
// 8, vectors as (re, im) pairs).
constexpr int kBlochLanes = 4;
// Blocks of a momentum-sector H.v launch, and of its dot partials.
inline unsigned bloch_blocks(const DevBloch& B) {
  if (!B.dk) return 1u;
  const uint64_t threads = (uint64_t)B.dk * (B.row_ptr ? kBlochLanes : 1);
  return (unsigned)((threads + kBlock - 1) / kBlock);
}
// y = H x in a momentum sector (ed_bloch.hip), x and y realified; with
// dot_out, per-block partials of Re <x, Hx> (bloch_blocks(B) of them).
template <class V>
hipError_t launch_bloch_hv(const DevLayout& L, const DevBloch& B, const V* x, V* y, double* dot_out,
                           hipStream_t st);

// The configuration-basis form (re + i im, D = the fixed-Sz sector's
// dimension) of a realified momentum-sector vector w (ed_bloch.hip).
template <class V>
hipError_t launch_bloch_expand(const DevBloch& B, uint32_t D, const V* w, double* re, double* im, hipStream_t st);

hipError_t launch_unrank(const DevLayout& L, uint64_t first, uint64_t count, uint32_t* out, hipStream_t st);
hipError_t launch_rank(const DevLayout& L, const uint32_t* states, uint64_t count, uint64_t* out,
                       hipStream_t st);
hipError_t launch_elements(const DevLayout& L, uint64_t row0, uint64_t nrows, int max_per_row,
                           uint32_t* cols, double* vals, int* counts, hipStream_t st);

}  // namespace alps_ed
