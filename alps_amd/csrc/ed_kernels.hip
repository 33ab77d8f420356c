// ed_kernels.hip -- gfx950 kernels of the exact-diagonalization hot path.
//
// H.v is matrix-free on the bit-encoded fixed-Sz basis (the product the
// reference computes as uBLAS prod over an explicit map-of-maps matrix,
// src/ietl/interface/ublas.h:57-63).  The rows of one sector split into
// segments that share the high word h of the state (see ed_layout.hpp); one
// workgroup owns one segment:
//   - the segment's slice of x is staged in LDS, so every low-low spin flip
//     is an LDS gather; its target row comes from a per-class table read
//     with coalesced 16-bit loads;
//   - h is uniform, so the high-high flips are scalar decisions whose x
//     reads are contiguous runs of a neighbouring segment (coalesced);
//   - a cross flip (one bit in h, one in the low word) reads one element of
//     a neighbouring segment, its rank coming from a per-class table;
//   - the diagonal is two table values plus the cross bonds.
// Every global load is a wave-uniform base plus a per-lane offset, so the
// address costs no vector ALU work.  Bond lists are padded with null bonds to
// the unrolled chunk sizes, so no load depends on a runtime bond count.
// Accumulation is FP64 whatever the vector storage.  Reductions are
// deterministic and atomic-free: producers write one partial per block,
// consumers add the partials in index order.
#include <hip/hip_runtime.h>

#include "ed_device.hpp"
#include "ed_kernels.hpp"

namespace alps_ed {

namespace {

using namespace dev;

// ---------------------------------------------------------------------------
// H.v: one workgroup per segment, longest segments first
// ---------------------------------------------------------------------------
template <class V, bool DOT, int W, int CLL, int CHH, int CX>
__global__ __launch_bounds__(kBlock, W) void hv_kernel(DevLayout L, const V* __restrict__ x, V* __restrict__ y,
                                                       double* __restrict__ dot_out) {
  hv_segment_at<V, DOT, CLL, CHH, CX>(L, ldc_struct(L.seg_rec + blockIdx.x), x, y,
                                      DOT ? dot_out + blockIdx.x : nullptr);
}

// seg_rec[i] for processing index i (the record hv_segment would look up).
__global__ __launch_bounds__(kBlock) void seg_rec_kernel(DevLayout L, SegRec* __restrict__ out) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= L.n_seg) return;
  out[i] = seg_record(L, L.proc_hi[i]);
}

template <class V, bool DOT, int W, int CLL, int CHH, int CX>
void launch_hv_with(const DevLayout& L, const V* x, V* y, double* dot_out, hipStream_t st, size_t lds) {
  hipLaunchKernelGGL((hv_kernel<V, DOT, W, CLL, CHH, CX>), dim3(hv_blocks(L)), dim3(kBlock), lds, st, L, x, y,
                     dot_out);
}

// The schedule table of ed_kernels.hpp as template arguments.
template <class V, bool DOT>
void launch_hv_schedule(const DevLayout& L, const V* x, V* y, double* dot_out, hipStream_t st, size_t lds,
                        int schedule) {
  constexpr int W = HvWaves<V>::value;
  if constexpr (sizeof(V) == 4) {
    static_assert(kHvSchedules == 6, "one case per schedule");
    switch (schedule) {
      case 1: launch_hv_with<V, DOT, 8, 4, 8, 2>(L, x, y, dot_out, st, lds); return;
      case 2: launch_hv_with<V, DOT, 6, 4, 8, 2>(L, x, y, dot_out, st, lds); return;
      case 3: launch_hv_with<V, DOT, 6, 4, 4, 2>(L, x, y, dot_out, st, lds); return;
      case 4: launch_hv_with<V, DOT, 8, 8, 4, 2>(L, x, y, dot_out, st, lds); return;
      case 5: launch_hv_with<V, DOT, 6, 8, 8, 4>(L, x, y, dot_out, st, lds); return;
      default: break;
    }
  }
  launch_hv_with<V, DOT, W, 4, 4, 2>(L, x, y, dot_out, st, lds);
}

// ---------------------------------------------------------------------------
// Per-high-word lists of a layout, built on the device (build_hv_lists)
// ---------------------------------------------------------------------------
// off[h + 1] = the high-high bonds that flip high word h (its two bits
// differ), rounded up to kHHList; 0 where h has no segment; off[0] = 0.
__global__ __launch_bounds__(kBlock) void hh_count_kernel(uint32_t nhi, const uint32_t* __restrict__ hi_base,
                                                          const uint32_t* __restrict__ mask, int nb,
                                                          uint32_t* __restrict__ off) {
  const uint32_t h = blockIdx.x * kBlock + threadIdx.x;
  if (h == 0) off[0] = 0u;
  if (h >= nhi) return;
  uint32_t n = 0;
  if (hi_base[h] != 0xffffffffu)
    for (int q = 0; q < nb; ++q) {
      const uint32_t bits = h & mask[q];
      n += (bits != 0u && bits != mask[q]) ? 1u : 0u;
    }
  off[h + 1] = (n + kHHList - 1) / kHHList * kHHList;
}

// In-place inclusive prefix sum of v[0..n) by one workgroup of kScanBlock
// threads: each thread sums a contiguous chunk, the chunk sums are scanned
// in LDS, then each thread rewrites its chunk.  Sums of 2^32 - 1 or more are
// stored as 0xffffffff (the caller rejects them).
constexpr int kScanBlock = 1024;
__global__ __launch_bounds__(kScanBlock) void scan_inclusive_kernel(uint32_t* __restrict__ v, uint32_t n) {
  __shared__ uint64_t part[kScanBlock];
  const uint32_t chunk = (n + kScanBlock - 1) / kScanBlock;
  const uint32_t b = threadIdx.x * chunk, e = min(n, b + chunk);
  uint64_t s = 0;
  for (uint32_t i = b; i < e; ++i) s += v[i];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int d = 1; d < kScanBlock; d <<= 1) {
    const uint64_t add = threadIdx.x >= (unsigned)d ? part[threadIdx.x - d] : 0u;
    __syncthreads();
    part[threadIdx.x] += add;
    __syncthreads();
  }
  uint64_t run = threadIdx.x > 0 ? part[threadIdx.x - 1] : 0u;
  for (uint32_t i = b; i < e; ++i) {
    run += v[i];
    v[i] = run >= 0xffffffffULL ? 0xffffffffu : (uint32_t)run;
  }
}

// The entries of high word h at [off[h], off[h+1]): its flipping high-high
// bonds in bond order (neighbour segment start, coefficient), then
// coefficient-0 entries that read h's own segment.
__global__ __launch_bounds__(kBlock) void hh_fill_kernel(uint32_t nhi, const uint32_t* __restrict__ hi_base,
                                                         const uint32_t* __restrict__ mask,
                                                         const double* __restrict__ t, int nb,
                                                         const uint32_t* __restrict__ off, HHEntry* __restrict__ ent) {
  const uint32_t h = blockIdx.x * kBlock + threadIdx.x;
  if (h >= nhi) return;
  const uint32_t own = hi_base[h];
  if (own == 0xffffffffu) return;
  uint32_t e = off[h];
  const uint32_t e1 = off[h + 1];
  for (int q = 0; q < nb; ++q) {
    const uint32_t bits = h & mask[q];
    if (bits != 0u && bits != mask[q]) ent[e++] = HHEntry{hi_base[h ^ mask[q]], 0u, t[q]};
  }
  for (; e < e1; ++e) ent[e] = HHEntry{own, 0u, 0.0};
}

// xb_nb[h * n_xb_p + c]: the first row of the segment of h ^ hbit[c], or of
// h's own when that high word has none; 0 where h has no segment.
__global__ __launch_bounds__(kBlock) void xb_nb_kernel(uint32_t nhi, int n_xb_p, const uint32_t* __restrict__ hi_base,
                                                       const uint32_t* __restrict__ hbit, uint32_t* __restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= (uint64_t)nhi * n_xb_p) return;
  const uint32_t h = (uint32_t)(i / n_xb_p), c = (uint32_t)(i % n_xb_p);
  const uint32_t own = hi_base[h];
  uint32_t v = 0u;
  if (own != 0xffffffffu) {
    const uint32_t nb = hi_base[h ^ hbit[c]];
    v = nb == 0xffffffffu ? own : nb;
  }
  out[i] = v;
}

// One workgroup per item; 16-byte moves where source and destination allow.
__global__ __launch_bounds__(kBlock) void scatter_copy_kernel(const unsigned char* __restrict__ src,
                                                              const ScatterItem* __restrict__ items) {
  const ScatterItem it = items[blockIdx.x];
  const unsigned char* s = src + it.src;
  unsigned char* d = static_cast<unsigned char*>(it.dst);
  const bool aligned = ((reinterpret_cast<uintptr_t>(s) | reinterpret_cast<uintptr_t>(d)) & 15u) == 0u;
  const uint64_t n16 = aligned ? it.bytes / 16u : 0u;
  for (uint64_t i = threadIdx.x; i < n16; i += kBlock)
    reinterpret_cast<uint4*>(d)[i] = reinterpret_cast<const uint4*>(s)[i];
  for (uint64_t i = n16 * 16u + threadIdx.x; i < it.bytes; i += kBlock) d[i] = s[i];
}

// ---------------------------------------------------------------------------
// Lanczos update, reductions, vector kernels
// ---------------------------------------------------------------------------
// `out` may alias w or u_prev: every element is read before the same thread
// writes it.
template <class V>
__global__ __launch_bounds__(kBlock) void update_kernel(uint64_t n, const V* w, const V* __restrict__ u_cur,
                                                        const V* u_prev, V* out, const double* __restrict__ alpha,
                                                        const double* __restrict__ s_cur,
                                                        const double* __restrict__ beta_prev,
                                                        const double* __restrict__ s_prev,
                                                        double* __restrict__ norm_out) {
  const double sc = s_cur ? *s_cur : 1.0;
  const double ca = (*alpha) * sc;
  const double cb = beta_prev ? (*beta_prev) * (s_prev ? *s_prev : 1.0) : 0.0;
  double acc = 0.0;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    double r = sc * (double)w[i] - ca * (double)u_cur[i];
    if (beta_prev) r -= cb * (double)u_prev[i];
    const V rs = (V)r;
    out[i] = rs;
    const double rd = (double)rs;
    acc += rd * rd;
  }
  store_partial(norm_out, acc);
}

// ---------------------------------------------------------------------------
// Partial reorthogonalization (ReorthDev, ed_kernels.hpp)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double block_max_all(double v) {
  __shared__ double max_smem[kBlock / 64];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  __syncthreads();
  if ((threadIdx.x & 63) == 0) max_smem[threadIdx.x >> 6] = v;
  __syncthreads();
  double m = max_smem[0];
#pragma unroll
  for (int w = 1; w < kBlock / 64; ++w) m = fmax(m, max_smem[w]);
  return m;
}

// The omega step of partial reorthogonalization for step j, computed by
// every block of reorth_dot_kernel with the same FP64 arithmetic on the same
// data (so every block reaches the same decision); block 0 stores the new
// row, the flag and the |T| estimate.  With the update of step j done
// (alpha_j, beta_{j-1} stored, r = u_{j+1} with norm partials norm_r), the
// omega recurrence gives row j+1 from rows j and j-1: for k < j
//   beta_j w_{j+1,k} = beta_k w_{j,k+1} + (alpha_k - alpha_j) w_{j,k}
//                      + beta_{k-1} w_{j,k-1} - beta_{j-1} w_{j-1,k} + theta,
// which follows from the Lanczos relation of steps j and k and the symmetry
// of H; theta bounds the rounding terms u_k.f_j - u_j.f_k by 2 eps |H| with
// the sign of the rest (a deterministic worst case), |H| estimated by the
// largest |alpha_i| + beta_i + beta_{i-1} so far.  w_{j+1,j} = 2 eps |H| /
// beta_j (the local loss), w_{j+1,j+1} = 1.  If the largest |w_{j+1,k}|
// exceeds sqrt(eps), or the previous step crossed it, the step is flagged
// and the row reset to eps (the Gram-Schmidt pass that follows makes
// u_{j+1} orthogonal to working accuracy).  Returns the flag.
__device__ int omega_step(int j, Partials norm_r, const ReorthDev& R) {
  const double* __restrict__ alpha = R.alpha;
  const double* __restrict__ beta = R.beta;
  const double aj = alpha[j];
  const double bjm1 = j > 0 ? beta[j - 1] : 0.0;
  const double bj = sqrt(sum_partials(norm_r));
  const double hn = fmax(j > 0 ? R.hnorm[j - 1] : 0.0, fabs(aj) + bj + bjm1);
  const double eps = R.eps;
  const double noise = 2.0 * eps * hn;
  const bool store = blockIdx.x == 0;
  double* w1 = R.omega + (size_t)((j + 1) % 3) * R.stride;
  const double* w0 = R.omega + (size_t)(j % 3) * R.stride;
  const double* wm = R.omega + (size_t)((j + 2) % 3) * R.stride;  // row j-1 (unused at j = 0)
  const bool ok = bj > 0.0 && isfinite(bj);
  double mx = 0.0;
  if (ok)
    for (int k = threadIdx.x; k < j; k += kBlock) {
      double t = beta[k] * w0[k + 1] + (alpha[k] - aj) * w0[k] - bjm1 * wm[k];
      if (k > 0) t += beta[k - 1] * w0[k - 1];
      const double v = (t + copysign(noise, t)) / bj;
      if (store) w1[k] = v;
      mx = fmax(mx, fabs(v));
    }
  mx = block_max_all(mx);
  const int prev = j > 0 ? R.flag[j - 1] : 0;
  const int flag = !ok ? 0 : (prev == 1 ? 2 : (mx > sqrt(eps) ? 1 : 0));
  if (store) {
    if (flag)
      for (int k = threadIdx.x; k < j; k += kBlock) w1[k] = eps;
    if (threadIdx.x == 0) {
      if (j == 0) R.omega[0] = 1.0;  // row 0: v_0.v_0
      w1[j] = ok ? noise / bj : eps;
      w1[j + 1] = 1.0;
      R.flag[j] = flag;
      R.hnorm[j] = hn;
    }
  }
  return flag;
}

// Partials of u_k . r for k = 0..j: block b owns a contiguous range of the
// four-element groups of the vectors and writes dot_part[k * G + b]; the
// vectors are read eight at a time with r (one read of r per eight vectors).
template <class V>
__global__ __launch_bounds__(kBlock) void reorth_dot_kernel(uint64_t n, int j, const V* __restrict__ r,
                                                            Partials norm_r, ReorthDev R) {
  if (omega_step(j, norm_r, R) == 0) return;
  typedef V v4 __attribute__((ext_vector_type(4), aligned(sizeof(V))));
  const V* const* slots = reinterpret_cast<const V* const*>(R.slots);
  const uint64_t n4 = n / 4;
  const uint64_t per = (n4 + gridDim.x - 1) / gridDim.x;
  const uint64_t q0 = (uint64_t)blockIdx.x * per;
  const uint64_t q1 = q0 + per < n4 ? q0 + per : n4;
  const int m = j + 1;
  constexpr int KG = 8;
  for (int k0 = 0; k0 < m; k0 += KG) {
    const V* u[KG];
#pragma unroll
    for (int c = 0; c < KG; ++c) u[c] = slots[k0 + c < m ? k0 + c : k0];
    double acc[KG];
#pragma unroll
    for (int c = 0; c < KG; ++c) acc[c] = 0.0;
    for (uint64_t q = q0 + threadIdx.x; q < q1; q += kBlock) {
      const v4 rv = *reinterpret_cast<const v4*>(r + 4 * q);
#pragma unroll
      for (int c = 0; c < KG; ++c) {
        if (k0 + c < m) {
          const v4 uv = *reinterpret_cast<const v4*>(u[c] + 4 * q);
#pragma unroll
          for (int e = 0; e < 4; ++e) acc[c] += (double)uv[e] * (double)rv[e];
        }
      }
    }
    // the last n mod 4 elements belong to block 0
    if (blockIdx.x == 0 && threadIdx.x < n - 4 * n4) {
      const uint64_t i = 4 * n4 + threadIdx.x;
#pragma unroll
      for (int c = 0; c < KG; ++c)
        if (k0 + c < m) acc[c] += (double)u[c][i] * (double)r[i];
    }
#pragma unroll
    for (int c = 0; c < KG; c += 2) {
      double a = acc[c], b = acc[c + 1];
      block_sum_all2(a, b);
      if (threadIdx.x == 0) {
        if (k0 + c < m) R.dot_part[(size_t)(k0 + c) * gridDim.x + blockIdx.x] = a;
        if (k0 + c + 1 < m) R.dot_part[(size_t)(k0 + c + 1) * gridDim.x + blockIdx.x] = b;
      }
    }
  }
}

// gamma_k = scale_k^2 * sum_b dot_part[k * G + b] (in block order), one wave
// per k.
__global__ __launch_bounds__(kBlock) void reorth_coef_kernel(int j, int G, const double* __restrict__ scale,
                                                             ReorthDev R) {
  if (R.flag[j] == 0) return;
  const int k = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (k > j) return;
  double s = 0.0;
  for (int b = lane; b < G; b += 64) s += R.dot_part[(size_t)k * G + b];
  s = wave_sum(s);
  if (lane == 0) R.gamma[k] = scale[k] * scale[k] * s;
}

// r -= sum_{k<=j} gamma_k u_k (FP64 per element, k ascending), stored in
// place; per-block partials of |r|^2 into norm_out (one per block of the
// grid, the same count as the update's).
template <class V>
__global__ __launch_bounds__(kBlock) void reorth_update_kernel(uint64_t n, int j, V* __restrict__ r, ReorthDev R,
                                                               double* __restrict__ norm_out) {
  if (R.flag[j] == 0) return;
  extern __shared__ double gam[];
  typedef V v4 __attribute__((ext_vector_type(4), aligned(sizeof(V))));
  const V* const* slots = reinterpret_cast<const V* const*>(R.slots);
  const int m = j + 1;
  for (int k = threadIdx.x; k < m; k += kBlock) gam[k] = R.gamma[k];
  __syncthreads();
  double acc = 0.0;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t n4 = n / 4;
  for (uint64_t q = gid; q < n4; q += stride) {
    const v4 rv = *reinterpret_cast<const v4*>(r + 4 * q);
    double o[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) o[e] = (double)rv[e];
    int k = 0;
    for (; k + 4 <= m; k += 4) {
      v4 uv[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) uv[c] = *reinterpret_cast<const v4*>(slots[k + c] + 4 * q);
#pragma unroll
      for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int e = 0; e < 4; ++e) o[e] -= gam[k + c] * (double)uv[c][e];
    }
    for (; k < m; ++k) {
      const v4 uv = *reinterpret_cast<const v4*>(slots[k] + 4 * q);
#pragma unroll
      for (int e = 0; e < 4; ++e) o[e] -= gam[k] * (double)uv[e];
    }
    v4 ov;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      ov[e] = (V)o[e];
      const double rd = (double)ov[e];
      acc += rd * rd;
    }
    *reinterpret_cast<v4*>(r + 4 * q) = ov;
  }
  for (uint64_t i = 4 * n4 + gid; i < n; i += stride) {
    double o = (double)r[i];
    for (int k = 0; k < m; ++k) o -= gam[k] * (double)slots[k][i];
    const V ov = (V)o;
    r[i] = ov;
    acc += (double)ov * (double)ov;
  }
  store_partial(norm_out, acc);
}

// One Lanczos update with its scalars: every block first forms, in the same
// order, scale_j = 1/sqrt(sum(norm_prev)) (the norm of u_j), beta_{j-1} =
// sqrt(sum(norm_prev)) when j > 0, and alpha_j = scale_j^2 * sum(hv_part)
// (both sums through one block reduction); block 0 stores them.  Then as
// update_kernel.  This replaces the two one-block reduction launches of a
// step.
template <class V>
__global__ __launch_bounds__(kBlock) void update_fused_kernel(uint64_t n, const V* w, const V* __restrict__ u_cur,
                                                              const V* u_prev, V* out, Partials hv_part,
                                                              Partials norm_prev, double* __restrict__ alpha_j,
                                                              double* __restrict__ beta_jm1,
                                                              double* __restrict__ scale_j,
                                                              const double* __restrict__ scale_jm1,
                                                              double* __restrict__ norm_out, double* alpha_h,
                                                              double* beta_h) {
  const bool has_prev = beta_jm1 != nullptr;
  const double s_prev = has_prev ? *scale_jm1 : 0.0;
  double nsum = thread_partials(norm_prev), hsum = thread_partials(hv_part);
  block_sum_all2(nsum, hsum);
  const double nrm = sqrt(nsum);
  const double sc = 1.0 / nrm;
  const double a = sc * sc * hsum;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    *alpha_j = a;
    *scale_j = sc;
    if (beta_jm1) *beta_jm1 = nrm;
    // host-mapped copies for the stopping test (no copy launch needed)
    if (alpha_h) *alpha_h = a;
    if (beta_h && beta_jm1) *beta_h = nrm;
  }
  const double ca = a * sc;
  const double cb = has_prev ? nrm * s_prev : 0.0;
  double acc = 0.0;
  // four consecutive elements per thread and iteration (vector loads)
  typedef V v4 __attribute__((ext_vector_type(4), aligned(sizeof(V))));
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t n4 = n / 4;
  for (uint64_t q = gid; q < n4; q += stride) {
    const v4 wv = *reinterpret_cast<const v4*>(w + 4 * q);
    const v4 uv = *reinterpret_cast<const v4*>(u_cur + 4 * q);
    const v4 pv = has_prev ? *reinterpret_cast<const v4*>(u_prev + 4 * q) : uv;
    v4 o;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      double r = sc * (double)wv[c] - ca * (double)uv[c];
      if (has_prev) r -= cb * (double)pv[c];
      o[c] = (V)r;
      const double rd = (double)o[c];
      acc += rd * rd;
    }
    *reinterpret_cast<v4*>(out + 4 * q) = o;
  }
  for (uint64_t i = 4 * n4 + gid; i < n; i += stride) {
    double r = sc * (double)w[i] - ca * (double)u_cur[i];
    if (has_prev) r -= cb * (double)u_prev[i];
    const V rs = (V)r;
    out[i] = rs;
    const double rd = (double)rs;
    acc += rd * rd;
  }
  store_partial(norm_out, acc);
}

template <class V>
__global__ __launch_bounds__(kBlock) void dot_kernel(uint64_t n, const V* __restrict__ x,
                                                     const V* __restrict__ y, double* partials) {
  double acc = 0.0;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    acc += (double)x[i] * (double)y[i];
  store_partial(partials, acc);
}

__global__ __launch_bounds__(kBlock) void finish_kernel(Partials in, const double* scale, double* out, double* out2,
                                                        int mode, double* out_h) {
  const double s = sum_partials(in);
  if (threadIdx.x == 0) {
    double v;
    if (mode == kNorm) {
      v = sqrt(s);
      if (out2) out2[0] = 1.0 / sqrt(s);
    } else if (mode == kAlpha) {
      const double sc = scale ? *scale : 1.0;
      v = sc * sc * s;
    } else {
      v = s;
    }
    out[0] = v;
    if (out_h) out_h[0] = v;
  }
}

template <class V>
__global__ __launch_bounds__(kBlock) void normalize_kernel(uint64_t n, V* v, Partials norm_in, double* b_out) {
  const double ss = sum_partials(norm_in);
  const double inv = 1.0 / sqrt(ss);
  if (blockIdx.x == 0 && threadIdx.x == 0 && b_out) *b_out = sqrt(ss);
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    v[i] = (V)(inv * (double)v[i]);
}

template <class V>
__global__ __launch_bounds__(kBlock) void fill_random_kernel(uint64_t n, V* v, uint64_t seed, double lo,
                                                             double hi) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    v[i] = (V)start_value(seed, i, lo, hi);
  }
}

template <class V>
__global__ __launch_bounds__(kBlock) void fill_kernel(uint64_t n, V* v, double value) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) v[i] = (V)value;
}

template <class V>
__global__ __launch_bounds__(kBlock) void axpby_kernel(uint64_t n, V* y, double a, const V* x, double b) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double xv = x ? (double)x[i] : 0.0;
    const double yv = b == 0.0 ? 0.0 : b * (double)y[i];
    y[i] = (V)(a * xv + yv);
  }
}

template <class VD, class VS>
__global__ __launch_bounds__(kBlock) void convert_kernel(uint64_t n, VD* __restrict__ dst,
                                                         const VS* __restrict__ src, const double* scale) {
  const double s = scale ? *scale : 1.0;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    dst[i] = (VD)(s * (double)src[i]);
}

// Ritz vectors over a stored Lanczos basis:
//   out_i[r] = sum_{j<m} (coef[i*m + j] * scale[j]) * u_j[r],  i < nvec <= NV,
// summed over j in ascending order.  Each thread takes four consecutive rows
// (vector loads) and four basis vectors at a time, so sixteen elements per
// thread are in flight.
template <class V, int NV>
__global__ __launch_bounds__(kBlock) void ritz_kernel(uint64_t n, const V* const* __restrict__ slots,
                                                      const double* __restrict__ scale, int m, int nvec,
                                                      const double* __restrict__ coef,
                                                      double* const* __restrict__ outs) {
  typedef V v4 __attribute__((ext_vector_type(4), aligned(sizeof(V))));
  typedef double d4 __attribute__((ext_vector_type(4), aligned(8)));
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t n4 = n / 4;
  for (uint64_t q = gid; q < n4; q += stride) {
    double acc[NV][4];
#pragma unroll
    for (int i = 0; i < NV; ++i)
#pragma unroll
      for (int e = 0; e < 4; ++e) acc[i][e] = 0.0;
    int j = 0;
    for (; j + 4 <= m; j += 4) {
      v4 u[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) u[c] = *reinterpret_cast<const v4*>(slots[j + c] + 4 * q);
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const double sj = scale[j + c];
#pragma unroll
        for (int i = 0; i < NV; ++i) {
          if (i < nvec) {
            const double w = coef[i * m + j + c] * sj;
#pragma unroll
            for (int e = 0; e < 4; ++e) acc[i][e] += w * (double)u[c][e];
          }
        }
      }
    }
    for (; j < m; ++j) {
      const v4 u = *reinterpret_cast<const v4*>(slots[j] + 4 * q);
      const double sj = scale[j];
#pragma unroll
      for (int i = 0; i < NV; ++i) {
        if (i < nvec) {
          const double w = coef[i * m + j] * sj;
#pragma unroll
          for (int e = 0; e < 4; ++e) acc[i][e] += w * (double)u[e];
        }
      }
    }
#pragma unroll
    for (int i = 0; i < NV; ++i)
      if (i < nvec) *reinterpret_cast<d4*>(outs[i] + 4 * q) = d4{acc[i][0], acc[i][1], acc[i][2], acc[i][3]};
  }
  for (uint64_t r = 4 * n4 + gid; r < n; r += stride) {
    double acc[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) acc[i] = 0.0;
    for (int j = 0; j < m; ++j) {
      const double u = (double)slots[j][r];
      const double sj = scale[j];
#pragma unroll
      for (int i = 0; i < NV; ++i)
        if (i < nvec) acc[i] += coef[i * m + j] * sj * u;
    }
#pragma unroll
    for (int i = 0; i < NV; ++i)
      if (i < nvec) outs[i][r] = acc[i];
  }
}

// ---------------------------------------------------------------------------
// Basis: unrank / rank / explicit elements
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void unrank_kernel(DevLayout L, uint64_t first, uint64_t count,
                                                        uint32_t* out) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < count) out[t] = unrank(L, (uint32_t)(first + t));
}

__global__ __launch_bounds__(kBlock) void rank_kernel(DevLayout L, const uint32_t* states, uint64_t count,
                                                      uint64_t* out) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < count) out[t] = rank_of(L, states[t]);
}

// Row elements in the reference's accumulation order
// (hamiltonian_matrix.hpp:44-58 explicit 0.0 diagonal; :269-291 all sites in
// order; :308-345 then every bond in edge order), columns ascending.
__global__ __launch_bounds__(64) void elements_kernel(DevLayout L, uint64_t row0, uint64_t nrows,
                                                      int max_per_row, uint32_t* cols, double* vals,
                                                      int* counts) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nrows) return;
  const uint32_t row = (uint32_t)(row0 + t);
  const uint32_t s = unrank(L, row);
  uint32_t* c = cols + t * (uint64_t)max_per_row;
  double* v = vals + t * (uint64_t)max_per_row;
  double d = 0.0;
  for (int site = 0; site < L.N; ++site) {
    const bool up = (s >> (L.N - 1 - site)) & 1u;
    const double e = up ? L.site_up[site] : L.site_dn[site];
    if (e != 0.0) d += e;
  }
  int n = 1;
  c[0] = row;
  v[0] = 0.0;
  for (int b = 0; b < L.n_bonds; ++b) {
    const uint32_t p1 = L.bond_p1[b], p2 = L.bond_p2[b];
    const bool par = ((s >> p1) & 1u) == ((s >> p2) & 1u);
    const double e = par ? L.bond_par[b] : L.bond_anti[b];
    if (e != 0.0) d += e;
    const double tb = L.bond_t[b];
    if (!par && tb != 0.0) {
      const uint32_t j = rank_of(L, s ^ ((1u << p1) | (1u << p2)));
      // keep columns ascending; equal columns accumulate in bond order
      int pos = 0;
      while (pos < n && c[pos] < j) ++pos;
      if (pos < n && c[pos] == j) {
        v[pos] += tb;
      } else if (n < max_per_row) {
        for (int q = n; q > pos; --q) {
          c[q] = c[q - 1];
          v[q] = v[q - 1];
        }
        c[pos] = j;
        v[pos] = 0.0 + tb;
        ++n;
      }
    }
  }
  for (int q = 0; q < n; ++q)
    if (c[q] == row) v[q] = d;
  counts[t] = n;
}

}  // namespace

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <class V>
hipError_t launch_hv(const DevLayout& L, const V* x, V* y, double* dot_out, hipStream_t st, int schedule) {
  if (L.n_seg == 0) return hipSuccess;
  if (schedule < 0 || schedule >= kHvSchedules || L.seg_rec == nullptr) return hipErrorInvalidValue;
  const size_t lds = (size_t)(L.max_seg + 1) * sizeof(double);  // the FP64 slice of hv_segment
  if (dot_out)
    launch_hv_schedule<V, true>(L, x, y, dot_out, st, lds, schedule);
  else
    launch_hv_schedule<V, false>(L, x, y, dot_out, st, lds, schedule);
  return hipGetLastError();
}

template <class V>
hipError_t launch_update(uint64_t n, const V* w, const V* u_cur, const V* u_prev, V* out, const double* alpha,
                         const double* s_cur, const double* beta_prev, const double* s_prev, double* norm_out,
                         hipStream_t st) {
  hipLaunchKernelGGL((update_kernel<V>), dim3(stream_blocks(n)), dim3(kBlock), 0, st, n, w, u_cur, u_prev, out,
                     alpha, s_cur, beta_prev, s_prev, norm_out);
  return hipGetLastError();
}

template <class V>
hipError_t launch_update_fused(uint64_t n, const V* w, const V* u_cur, const V* u_prev, V* out, Partials hv_part,
                               Partials norm_prev, double* alpha_j, double* beta_jm1, double* scale_j,
                               const double* scale_jm1, double* norm_out, hipStream_t st, double* alpha_h,
                               double* beta_h) {
  hipLaunchKernelGGL((update_fused_kernel<V>), dim3(fused_blocks(n)), dim3(kBlock), 0, st, n, w, u_cur, u_prev, out,
                     hv_part, norm_prev, alpha_j, beta_jm1, scale_j, scale_jm1, norm_out, alpha_h, beta_h);
  return hipGetLastError();
}

template <class V>
hipError_t launch_reorth(uint64_t n, int j, V* r, const double* scale, const ReorthDev& R, double* norm_out,
                         hipStream_t st) {
  if (j < 0 || (uint32_t)j + 2 > R.stride) return hipErrorInvalidValue;
  const int G = kReorthDotBlocks;
  hipLaunchKernelGGL((reorth_dot_kernel<V>), dim3(G), dim3(kBlock), 0, st, n, j, r,
                     Partials{norm_out, fused_blocks(n)}, R);
  const unsigned cb = (unsigned)((j + 1 + kBlock / 64 - 1) / (kBlock / 64));
  hipLaunchKernelGGL(reorth_coef_kernel, dim3(cb), dim3(kBlock), 0, st, j, G, scale, R);
  hipLaunchKernelGGL((reorth_update_kernel<V>), dim3(fused_blocks(n)), dim3(kBlock), (size_t)(j + 1) * sizeof(double),
                     st, n, j, r, R, norm_out);
  return hipGetLastError();
}

template <class V>
hipError_t launch_dot(uint64_t n, const V* x, const V* y, double* partials, hipStream_t st) {
  hipLaunchKernelGGL((dot_kernel<V>), dim3(stream_blocks(n)), dim3(kBlock), 0, st, n, x, y, partials);
  return hipGetLastError();
}

hipError_t launch_finish(Partials in, const double* scale, double* out, double* out2, int mode, hipStream_t st,
                         double* out_h) {
  hipLaunchKernelGGL(finish_kernel, dim3(1), dim3(kBlock), 0, st, in, scale, out, out2, mode, out_h);
  return hipGetLastError();
}

template <class V>
hipError_t launch_normalize(uint64_t n, V* v, Partials norm_in, double* b_out, hipStream_t st) {
  hipLaunchKernelGGL((normalize_kernel<V>), dim3(stream_blocks(n)), dim3(kBlock), 0, st, n, v, norm_in, b_out);
  return hipGetLastError();
}

hipError_t launch_hh_count(uint32_t nhi, const uint32_t* hi_base, const uint32_t* mask, int nb, uint32_t* off,
                           const uint32_t* hbit, int n_xb_p, uint32_t* xb_nb, hipStream_t st) {
  const unsigned g = (unsigned)((nhi + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(hh_count_kernel, dim3(g), dim3(kBlock), 0, st, nhi, hi_base, mask, nb, off);
  hipLaunchKernelGGL(scan_inclusive_kernel, dim3(1), dim3(kScanBlock), 0, st, off + 1, nhi);
  const uint64_t nx = (uint64_t)nhi * n_xb_p;
  if (nx > 0)
    hipLaunchKernelGGL(xb_nb_kernel, dim3((unsigned)((nx + kBlock - 1) / kBlock)), dim3(kBlock), 0, st, nhi, n_xb_p,
                       hi_base, hbit, xb_nb);
  return hipGetLastError();
}

hipError_t launch_seg_rec(const DevLayout& L, SegRec* out, hipStream_t st) {
  if (L.n_seg == 0) return hipSuccess;
  hipLaunchKernelGGL(seg_rec_kernel, dim3((L.n_seg + kBlock - 1) / kBlock), dim3(kBlock), 0, st, L, out);
  return hipGetLastError();
}

hipError_t launch_hh_fill(uint32_t nhi, const uint32_t* hi_base, const uint32_t* mask, const double* t, int nb,
                          const uint32_t* off, HHEntry* ent, hipStream_t st) {
  const unsigned g = (unsigned)((nhi + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(hh_fill_kernel, dim3(g), dim3(kBlock), 0, st, nhi, hi_base, mask, t, nb, off, ent);
  return hipGetLastError();
}

hipError_t launch_scatter_copy(const unsigned char* src, const ScatterItem* items, uint32_t n, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(scatter_copy_kernel, dim3(n), dim3(kBlock), 0, st, src, items);
  return hipGetLastError();
}

template <class V>
hipError_t launch_fill_random(uint64_t n, V* v, uint64_t seed, double lo, double hi, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL((fill_random_kernel<V>), dim3(stream_blocks(n)), dim3(kBlock), 0, st, n, v, seed, lo, hi);
  return hipGetLastError();
}

template <class V>
hipError_t launch_fill(uint64_t n, V* v, double value, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL((fill_kernel<V>), dim3(stream_blocks(n)), dim3(kBlock), 0, st, n, v, value);
  return hipGetLastError();
}

template <class V>
hipError_t launch_axpby(uint64_t n, V* y, double a, const V* x, double b, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL((axpby_kernel<V>), dim3(stream_blocks(n)), dim3(kBlock), 0, st, n, y, a, x, b);
  return hipGetLastError();
}

template <class VD, class VS>
hipError_t launch_convert(uint64_t n, VD* dst, const VS* src, const double* scale, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL((convert_kernel<VD, VS>), dim3(stream_blocks(n)), dim3(kBlock), 0, st, n, dst, src, scale);
  return hipGetLastError();
}

template <class V>
hipError_t launch_ritz(uint64_t n, const V* const* slots, const double* scale, int m, int nvec,
                       const double* coef, double* const* outs, hipStream_t st) {
  if (n == 0 || nvec == 0) return hipSuccess;
  if (nvec < 0 || nvec > kMaxRitz) return hipErrorInvalidValue;
  const dim3 g(stream_blocks(n)), b(kBlock);
  if (nvec == 1)
    hipLaunchKernelGGL((ritz_kernel<V, 1>), g, b, 0, st, n, slots, scale, m, nvec, coef, outs);
  else if (nvec == 2)
    hipLaunchKernelGGL((ritz_kernel<V, 2>), g, b, 0, st, n, slots, scale, m, nvec, coef, outs);
  else if (nvec <= 4)
    hipLaunchKernelGGL((ritz_kernel<V, 4>), g, b, 0, st, n, slots, scale, m, nvec, coef, outs);
  else
    hipLaunchKernelGGL((ritz_kernel<V, kMaxRitz>), g, b, 0, st, n, slots, scale, m, nvec, coef, outs);
  return hipGetLastError();
}

hipError_t launch_unrank(const DevLayout& L, uint64_t first, uint64_t count, uint32_t* out, hipStream_t st) {
  if (count == 0) return hipSuccess;
  const unsigned nb = (unsigned)((count + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(unrank_kernel, dim3(nb), dim3(kBlock), 0, st, L, first, count, out);
  return hipGetLastError();
}

hipError_t launch_rank(const DevLayout& L, const uint32_t* states, uint64_t count, uint64_t* out,
                       hipStream_t st) {
  if (count == 0) return hipSuccess;
  const unsigned nb = (unsigned)((count + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(rank_kernel, dim3(nb), dim3(kBlock), 0, st, L, states, count, out);
  return hipGetLastError();
}

hipError_t launch_elements(const DevLayout& L, uint64_t row0, uint64_t nrows, int max_per_row,
                           uint32_t* cols, double* vals, int* counts, hipStream_t st) {
  if (nrows == 0) return hipSuccess;
  const unsigned nb = (unsigned)((nrows + 63) / 64);
  hipLaunchKernelGGL(elements_kernel, dim3(nb), dim3(64), 0, st, L, row0, nrows, max_per_row, cols, vals,
                     counts);
  return hipGetLastError();
}

#define ALPS_ED_INSTANTIATE(V)                                                                              \
  template hipError_t launch_hv<V>(const DevLayout&, const V*, V*, double*, hipStream_t, int);              \
  template hipError_t launch_update<V>(uint64_t, const V*, const V*, const V*, V*, const double*,           \
                                       const double*, const double*, const double*, double*, hipStream_t); \
  template hipError_t launch_dot<V>(uint64_t, const V*, const V*, double*, hipStream_t);                    \
  template hipError_t launch_update_fused<V>(uint64_t, const V*, const V*, const V*, V*, Partials, Partials,   \
                                             double*, double*, double*, const double*, double*, hipStream_t,   \
                                             double*, double*);                                                \
  template hipError_t launch_normalize<V>(uint64_t, V*, Partials, double*, hipStream_t);                    \
  template hipError_t launch_fill_random<V>(uint64_t, V*, uint64_t, double, double, hipStream_t);           \
  template hipError_t launch_fill<V>(uint64_t, V*, double, hipStream_t);                                    \
  template hipError_t launch_axpby<V>(uint64_t, V*, double, const V*, double, hipStream_t);                 \
  template hipError_t launch_ritz<V>(uint64_t, const V* const*, const double*, int, int, const double*,    \
                                     double* const*, hipStream_t);                                          \
  template hipError_t launch_reorth<V>(uint64_t, int, V*, const double*, const ReorthDev&, double*, hipStream_t);

ALPS_ED_INSTANTIATE(float)
ALPS_ED_INSTANTIATE(double)
template hipError_t launch_convert<float, float>(uint64_t, float*, const float*, const double*, hipStream_t);
template hipError_t launch_convert<float, double>(uint64_t, float*, const double*, const double*, hipStream_t);
template hipError_t launch_convert<double, float>(uint64_t, double*, const float*, const double*, hipStream_t);
template hipError_t launch_convert<double, double>(uint64_t, double*, const double*, const double*, hipStream_t);

}  // namespace alps_ed
