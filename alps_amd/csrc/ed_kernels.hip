// ed_kernels.hip -- gfx950 kernels of the exact-diagonalization hot path.
//
// H.v is matrix-free on the bit-encoded fixed-Sz basis (the product the
// reference computes as uBLAS prod over an explicit map-of-maps matrix,
// src/ietl/interface/ublas.h:57-63).  One wavefront owns a 64-row tile of a
// row segment that shares one high word `h`; lanes run over consecutive low
// words.  Consequences on CDNA4:
//   - h is wave-uniform (readfirstlane), so everything that depends on it --
//     the high-word diagonal, the high-high flip bonds and their target
//     segments -- is scalar work, and the x loads of high-high flips are
//     64-lane contiguous runs (fully coalesced);
//   - low-low flips stay inside the segment (a window of at most C(B,B/2)
//     rows), cross flips land in one neighbouring segment;
//   - the diagonal comes from two table lookups plus the cross bonds.
// Accumulation is FP64 whatever the vector storage; reductions are
// deterministic (fixed tree + last-block finish, no float atomics).
#include <hip/hip_runtime.h>

#include "ed_kernels.hpp"

namespace alps_ed {

namespace {

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Block sum, valid in thread 0.
__device__ __forceinline__ double block_sum(double v) {
  __shared__ double red_smem[kWavesPerBlock];
  v = wave_sum(v);
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) red_smem[wave] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int w = 0; w < kWavesPerBlock; ++w) s += red_smem[w];
  }
  __syncthreads();
  return s;
}

__device__ __forceinline__ double load_coherent(const double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Grid reduction finish: called by every thread of every block with the
// thread's partial value.
__device__ void grid_reduce(double v, const GridReduce& r) {
  __shared__ bool is_last;
  const double bs = block_sum(v);
  if (threadIdx.x == 0) {
    r.partials[blockIdx.x] = bs;
    __threadfence();
    const unsigned done = atomicAdd(r.counter, 1u);
    is_last = (done == gridDim.x - 1);
  }
  __syncthreads();
  if (!is_last) return;
  __threadfence();
  double s = 0.0;
  for (unsigned i = threadIdx.x; i < gridDim.x; i += blockDim.x) s += load_coherent(r.partials + i);
  s = block_sum(s);
  if (threadIdx.x == 0) {
    if (r.mode == 0) {
      r.out[0] = s;
    } else {
      const double nrm = sqrt(s);
      r.out[0] = nrm;
      if (r.out2) r.out2[0] = 1.0 / nrm;
    }
    *r.counter = 0u;
  }
}

template <class V>
__device__ __forceinline__ double ld(const V* p, uint32_t i) {
  return (double)p[i];
}

// ---------------------------------------------------------------------------
// H.v
// ---------------------------------------------------------------------------
template <class V, bool DOT>
__global__ __launch_bounds__(kBlock) void hv_kernel(DevLayout L, const V* __restrict__ x,
                                                     V* __restrict__ y, const double* xscale,
                                                     GridReduce red) {
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t tile = blockIdx.x * kWavesPerBlock + wave;
  const double s = xscale ? *xscale : 1.0;
  double dotp = 0.0;
  if (tile < L.n_tiles) {
    const uint32_t h = __builtin_amdgcn_readfirstlane(L.tiles[2 * tile]);
    const uint32_t r0 = __builtin_amdgcn_readfirstlane(L.tiles[2 * tile + 1]);
    const int kl = L.k - __builtin_popcount(h);
    const uint32_t cls0 = L.lo_class_off[kl];
    const uint32_t len = L.lo_class_off[kl + 1] - cls0;
    const uint32_t base = L.hi_base[h];
    const double dh = L.diag_hi[h];
    const uint32_t r = r0 + lane;
    if (r < len) {
      const uint32_t lo = L.lo_state[cls0 + r];
      const uint32_t row = base + r;
      const double xi = ld(x, row);
      double d = dh + L.diag_lo[lo];
      double off = 0.0;
      // cross bonds: one bit in h (uniform), one in lo
      for (int b = 0; b < L.n_xb; ++b) {
        const CrossBond cb = L.xb[b];
        const bool hb = (h & cb.hbit) != 0u;
        const bool lb = (lo & cb.lbit) != 0u;
        if (hb == lb) {
          d += cb.e_par;
        } else {
          d += cb.e_anti;
          const uint32_t j = L.hi_base[h ^ cb.hbit] + L.lo_rank[lo ^ cb.lbit];
          off += cb.t * ld(x, j);
        }
      }
      // low-low flips: target in the same segment
      for (int b = 0; b < L.n_ll; ++b) {
        const uint32_t m = L.ll[b].mask;
        const uint32_t bits = lo & m;
        if (bits != 0u && bits != m) off += L.ll[b].t * ld(x, base + L.lo_rank[lo ^ m]);
      }
      // high-high flips: wave-uniform test, coalesced 64-row target run
      for (int b = 0; b < L.n_hh; ++b) {
        const uint32_t m = L.hh[b].mask;
        const uint32_t bits = h & m;
        if (bits != 0u && bits != m) off += L.hh[b].t * ld(x, L.hi_base[h ^ m] + r);
      }
      const double yv = s * (d * xi + off);
      const V ys = (V)yv;
      y[row] = ys;
      if (DOT) dotp = (s * xi) * (double)ys;
    }
  }
  if (DOT) grid_reduce(dotp, red);
}

// ---------------------------------------------------------------------------
// Lanczos update + norm
// ---------------------------------------------------------------------------
// `out` may alias w or u_prev: every element is read before it is written by
// the same thread.
template <class V>
__global__ __launch_bounds__(kBlock) void update_kernel(uint64_t n, const V* w,
                                                        const V* __restrict__ u_cur,
                                                        const V* u_prev, V* out, const double* alpha,
                                                        const double* s_cur, const double* beta_prev,
                                                        const double* s_prev, GridReduce red) {
  const double ca = (*alpha) * (s_cur ? *s_cur : 1.0);
  const double cb = beta_prev ? (*beta_prev) * (s_prev ? *s_prev : 1.0) : 0.0;
  double acc = 0.0;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    double r = (double)w[i] - ca * (double)u_cur[i];
    if (beta_prev) r -= cb * (double)u_prev[i];
    const V rs = (V)r;
    out[i] = rs;
    const double rd = (double)rs;
    acc += rd * rd;
  }
  grid_reduce(acc, red);
}

template <class V>
__global__ __launch_bounds__(kBlock) void dot_kernel(uint64_t n, const V* __restrict__ x,
                                                     const V* __restrict__ y, GridReduce red) {
  double acc = 0.0;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    acc += (double)x[i] * (double)y[i];
  grid_reduce(acc, red);
}

__device__ __forceinline__ uint64_t splitmix64_mix(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

template <class V>
__global__ __launch_bounds__(kBlock) void fill_random_kernel(uint64_t n, V* v, uint64_t seed,
                                                             double lo, double hi) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint64_t z = splitmix64_mix(seed + (i + 1) * 0x9e3779b97f4a7c15ULL);
    const double u = (double)(z >> 11) * (1.0 / 9007199254740992.0);
    v[i] = (V)(lo + (hi - lo) * u);
  }
}

template <class V>
__global__ __launch_bounds__(kBlock) void fill_kernel(uint64_t n, V* v, double value) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) v[i] = (V)value;
}

template <class V>
__global__ __launch_bounds__(kBlock) void axpby_kernel(uint64_t n, V* __restrict__ y, double a,
                                                       const V* __restrict__ x, double b,
                                                       const double* ap, const double* bp,
                                                       bool a_recip) {
  double av = ap ? *ap : a;
  if (a_recip) av = 1.0 / av;
  const double bv = bp ? *bp : b;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double xv = x ? (double)x[i] : 0.0;
    const double yv = bv == 0.0 ? 0.0 : bv * (double)y[i];
    y[i] = (V)(av * xv + yv);
  }
}

template <class VD, class VS>
__global__ __launch_bounds__(kBlock) void convert_kernel(uint64_t n, VD* __restrict__ dst,
                                                         const VS* __restrict__ src,
                                                         const double* scale) {
  const double s = scale ? *scale : 1.0;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    dst[i] = (VD)(s * (double)src[i]);
}

// ---------------------------------------------------------------------------
// Basis: unrank / rank / explicit elements
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t unrank(const DevLayout& L, uint32_t i) {
  uint32_t lo = 0, hi = L.n_seg - 1;
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) >> 1;
    if (L.seg_start[mid] <= i) lo = mid;
    else hi = mid - 1;
  }
  const uint32_t h = L.seg_hi[lo];
  const uint32_t r = i - L.seg_start[lo];
  const int kl = L.k - __builtin_popcount(h);
  return (h << L.B) | L.lo_state[L.lo_class_off[kl] + r];
}

__device__ __forceinline__ uint32_t rank_of(const DevLayout& L, uint32_t s) {
  if (__builtin_popcount(s) != L.k) return L.D;
  if (L.N < 32 && (s >> L.N) != 0u) return L.D;
  const uint32_t h = L.NH > 0 ? (s >> L.B) : 0u;
  const uint32_t lo = L.B >= 32 ? s : (s & ((1u << L.B) - 1u));
  const uint32_t hb = L.hi_base[h];
  if (hb == 0xffffffffu) return L.D;
  return hb + L.lo_rank[lo];
}

__global__ __launch_bounds__(kBlock) void unrank_kernel(DevLayout L, uint64_t first, uint64_t count,
                                                        uint32_t* out) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < count) out[t] = unrank(L, (uint32_t)(first + t));
}

__global__ __launch_bounds__(kBlock) void rank_kernel(DevLayout L, const uint32_t* states,
                                                      uint64_t count, uint64_t* out) {
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
  for (int b = 0; b < L.n_bonds; ++b) {
    const uint32_t p1 = L.bond_p1[b], p2 = L.bond_p2[b];
    const bool par = ((s >> p1) & 1u) == ((s >> p2) & 1u);
    const double e = par ? L.bond_par[b] : L.bond_anti[b];
    if (e != 0.0) d += e;
    const double tb = L.bond_t[b];
    if (!par && tb != 0.0) {
      const uint32_t j = rank_of(L, s ^ ((1u << p1) | (1u << p2)));
      // insert keeping columns ascending; equal columns accumulate in bond order
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
  // the diagonal slot was inserted first with a placeholder: fill it now
  for (int q = 0; q < n; ++q)
    if (c[q] == row) v[q] = d;
  counts[t] = n;
}

}  // namespace

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <class V>
hipError_t launch_hv(const DevLayout& L, const V* x, V* y, const double* xscale, GridReduce red,
                     hipStream_t st) {
  if (L.n_tiles == 0) return hipSuccess;
  const unsigned nb = hv_blocks(L);
  if (red.out)
    hipLaunchKernelGGL((hv_kernel<V, true>), dim3(nb), dim3(kBlock), 0, st, L, x, y, xscale, red);
  else
    hipLaunchKernelGGL((hv_kernel<V, false>), dim3(nb), dim3(kBlock), 0, st, L, x, y, xscale, red);
  return hipGetLastError();
}

template <class V>
hipError_t launch_update(uint64_t n, const V* w, const V* u_cur, const V* u_prev, V* out,
                         const double* alpha, const double* s_cur, const double* beta_prev,
                         const double* s_prev, GridReduce red, hipStream_t st) {
  hipLaunchKernelGGL((update_kernel<V>), dim3(stream_blocks(n)), dim3(kBlock), 0, st, n, w, u_cur,
                     u_prev, out, alpha, s_cur, beta_prev, s_prev, red);
  return hipGetLastError();
}

template <class V>
hipError_t launch_dot(uint64_t n, const V* x, const V* y, GridReduce red, hipStream_t st) {
  hipLaunchKernelGGL((dot_kernel<V>), dim3(stream_blocks(n)), dim3(kBlock), 0, st, n, x, y, red);
  return hipGetLastError();
}

template <class V>
hipError_t launch_fill_random(uint64_t n, V* v, uint64_t seed, double lo, double hi, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL((fill_random_kernel<V>), dim3(stream_blocks(n)), dim3(kBlock), 0, st, n, v, seed,
                     lo, hi);
  return hipGetLastError();
}

template <class V>
hipError_t launch_fill(uint64_t n, V* v, double value, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL((fill_kernel<V>), dim3(stream_blocks(n)), dim3(kBlock), 0, st, n, v, value);
  return hipGetLastError();
}

template <class V>
hipError_t launch_axpby(uint64_t n, V* y, double a, const V* x, double b, const double* ap,
                        const double* bp, bool a_recip, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL((axpby_kernel<V>), dim3(stream_blocks(n)), dim3(kBlock), 0, st, n, y, a, x, b,
                     ap, bp, a_recip);
  return hipGetLastError();
}

template <class VD, class VS>
hipError_t launch_convert(uint64_t n, VD* dst, const VS* src, const double* scale, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL((convert_kernel<VD, VS>), dim3(stream_blocks(n)), dim3(kBlock), 0, st, n, dst,
                     src, scale);
  return hipGetLastError();
}

hipError_t launch_unrank(const DevLayout& L, uint64_t first, uint64_t count, uint32_t* out,
                         hipStream_t st) {
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
  hipLaunchKernelGGL(elements_kernel, dim3(nb), dim3(64), 0, st, L, row0, nrows, max_per_row, cols,
                     vals, counts);
  return hipGetLastError();
}

#define ALPS_ED_INSTANTIATE(V)                                                                     \
  template hipError_t launch_hv<V>(const DevLayout&, const V*, V*, const double*, GridReduce,      \
                                   hipStream_t);                                                    \
  template hipError_t launch_update<V>(uint64_t, const V*, const V*, const V*, V*, const double*, \
                                       const double*, const double*, const double*, GridReduce,    \
                                       hipStream_t);                                                \
  template hipError_t launch_dot<V>(uint64_t, const V*, const V*, GridReduce, hipStream_t);        \
  template hipError_t launch_fill_random<V>(uint64_t, V*, uint64_t, double, double, hipStream_t);  \
  template hipError_t launch_fill<V>(uint64_t, V*, double, hipStream_t);                           \
  template hipError_t launch_axpby<V>(uint64_t, V*, double, const V*, double, const double*,       \
                                      const double*, bool, hipStream_t);

ALPS_ED_INSTANTIATE(float)
ALPS_ED_INSTANTIATE(double)
template hipError_t launch_convert<float, float>(uint64_t, float*, const float*, const double*, hipStream_t);
template hipError_t launch_convert<float, double>(uint64_t, float*, const double*, const double*, hipStream_t);
template hipError_t launch_convert<double, float>(uint64_t, double*, const float*, const double*, hipStream_t);
template hipError_t launch_convert<double, double>(uint64_t, double*, const double*, const double*, hipStream_t);

}  // namespace alps_ed
