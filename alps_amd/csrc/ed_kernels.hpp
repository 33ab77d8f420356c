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
  uint32_t n_seg, n_tiles;
  int n_ll, n_hh, n_xb, n_bonds;
  const uint32_t* lo_rank;
  const uint32_t* lo_state;
  const uint32_t* lo_class_off;
  const uint32_t* hi_base;
  const uint32_t* seg_hi;
  const uint32_t* seg_start;
  const uint32_t* tiles;
  const double* diag_lo;
  const double* diag_hi;
  const FlipBond* ll;
  const FlipBond* hh;
  const CrossBond* xb;
  const double* site_dn;
  const double* site_up;
  const uint32_t* bond_p1;
  const uint32_t* bond_p2;
  const double* bond_par;
  const double* bond_anti;
  const double* bond_t;
};

// A deterministic grid reduction: every block stores its partial sum, the
// last block to finish (counted with a device-scope atomic) adds the
// partials in index order and writes the result.
//   mode 0: out[0] = sum
//   mode 1: out[0] = sqrt(sum), out2[0] = 1/sqrt(sum)
struct GridReduce {
  double* partials;   // >= gridDim.x entries
  unsigned* counter;  // zero before the launch; reset to zero by the last block
  double* out;
  double* out2;
  int mode;
};

constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / 64;

// Number of blocks of the streaming kernels for a vector of length n.
inline unsigned stream_blocks(uint64_t n) {
  uint64_t b = (n + (uint64_t)kBlock * 4 - 1) / ((uint64_t)kBlock * 4);
  if (b < 1) b = 1;
  if (b > 2048) b = 2048;
  return (unsigned)b;
}

inline unsigned hv_blocks(const DevLayout& L) {
  unsigned b = (L.n_tiles + kWavesPerBlock - 1) / kWavesPerBlock;
  return b ? b : 1;
}

// y = s * H x (s = *xscale when xscale != nullptr, else 1).  With red.out
// set, also reduces alpha = sum_i (s x_i) * y_i (y as stored).
template <class V>
hipError_t launch_hv(const DevLayout& L, const V* x, V* y, const double* xscale, GridReduce red,
                     hipStream_t st);

// Lanczos update: r = w - (alpha*s_cur) u_cur - (beta_prev*s_prev) u_prev,
// stored into `out` (which may alias w or u_prev); reduces |r| into red
// (mode 1).  Scalars are read from device memory; beta_prev == nullptr drops
// the last term (first step); null s_cur / s_prev mean 1.
template <class V>
hipError_t launch_update(uint64_t n, const V* w, const V* u_cur, const V* u_prev, V* out,
                         const double* alpha, const double* s_cur, const double* beta_prev,
                         const double* s_prev, GridReduce red, hipStream_t st);

template <class V>
hipError_t launch_dot(uint64_t n, const V* x, const V* y, GridReduce red, hipStream_t st);
template <class V>
hipError_t launch_fill_random(uint64_t n, V* v, uint64_t seed, double lo, double hi, hipStream_t st);
template <class V>
hipError_t launch_fill(uint64_t n, V* v, double value, hipStream_t st);
// y = a*x + b*y; a and b from device memory when ap/bp are set.
template <class V>
hipError_t launch_axpby(uint64_t n, V* y, double a, const V* x, double b, const double* ap,
                        const double* bp, bool a_recip, hipStream_t st);
template <class VD, class VS>
hipError_t launch_convert(uint64_t n, VD* dst, const VS* src, const double* scale, hipStream_t st);

hipError_t launch_unrank(const DevLayout& L, uint64_t first, uint64_t count, uint32_t* out, hipStream_t st);
hipError_t launch_rank(const DevLayout& L, const uint32_t* states, uint64_t count, uint64_t* out,
                       hipStream_t st);
hipError_t launch_elements(const DevLayout& L, uint64_t row0, uint64_t nrows, int max_per_row,
                           uint32_t* cols, double* vals, int* counts, hipStream_t st);

}  // namespace alps_ed
