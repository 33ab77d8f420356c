// ed_measure.hip -- expectation values of spin operators in a vector of a
// fixed-Sz sector, for the eigenvector measurements of
// DiagMatrix::perform_measurements (applications/diag/diag.h:191-314), which
// evaluates <v| O |v> (SparseDiagMatrix::calculate, sparsediag.h:151-159)
// for site terms, site-pair terms and bond terms built from Sz, S+ and S-.
//
// Every such term is a combination of two families of sums over the rows r
// of the sector (w_r = v_r^2, s_i(r) = +-1/2 the Sz of site i in row r):
//   - diagonal moments: sum w, sum w s_i, sum w s_i s_j for all sites;
//   - flip amplitudes F(i,j) = sum over rows with site j up and site i down
//     of v_r v_r', r' the row with both spins turned (<v|S+_i S-_j|v>).
//
// Diagonal moments in two passes: one workgroup per row segment (shared
// high word h) reduces the segment's rows to the low-word statistics
// W = sum w, M_a = sum w s_a and P_ab = sum w s_a s_b over the B low bits
// (1 + B + B(B-1)/2 values); then one thread per output entry combines them
// over the segments in index order with the high bits of h, which are
// constant in a segment.  Flip amplitudes: workgroups stride over the
// segments, eight site pairs per workgroup row of the grid, target rows by
// the split ranking; one partial per workgroup, summed in index order.
// All sums are FP64 and deterministic.
#include <hip/hip_runtime.h>

#include "ed_device.hpp"
#include "ed_kernels.hpp"

namespace alps_ed {

namespace {

using namespace dev;

constexpr int kPairChunk = 8;  // site pairs per workgroup of the flip kernel

__device__ __forceinline__ double half_spin(uint32_t word, int bit) { return ((word >> bit) & 1u) ? 0.5 : -0.5; }

// Position of the low-word pair (a, b), a < b, in the order a = 0.., b = a+1..
__device__ __forceinline__ int lo_pair_index(int a, int b, int B) { return a * (2 * B - a - 1) / 2 + (b - a - 1); }

template <class V>
__global__ __launch_bounds__(kBlock) void moment_stats_kernel(DevLayout L, const V* __restrict__ v,
                                                              double* __restrict__ stats, int n_stats) {
  extern __shared__ __align__(16) unsigned char smem[];
  double* w = reinterpret_cast<double*>(smem);
  uint32_t* lo = reinterpret_cast<uint32_t*>(w + L.max_seg);
  const uint32_t h = L.seg_hi[blockIdx.x];
  const int kl = L.k - __builtin_popcount(h);
  const uint32_t cls0 = L.lo_class_off[kl];
  const uint32_t len = L.lo_class_off[kl + 1] - cls0;
  const uint32_t base = L.hi_base[h];
  for (uint32_t r = threadIdx.x; r < len; r += kBlock) {
    const double x = (double)v[base + r];
    w[r] = x * x;
    lo[r] = L.lo_state[cls0 + r];
  }
  __syncthreads();
  const int q = (int)threadIdx.x;
  if (q >= n_stats) return;
  const int B = L.B;
  int a = -1, b = -1;  // q = 0: W; a only: M_a; a and b: P_ab
  if (q >= 1 && q <= B) {
    a = q - 1;
  } else if (q > B) {
    int rest = q - 1 - B;
    for (a = 0; a < B; ++a) {
      const int cnt = B - a - 1;
      if (rest < cnt) break;
      rest -= cnt;
    }
    b = a + 1 + rest;
  }
  double acc = 0.0;
  if (a < 0) {
    for (uint32_t r = 0; r < len; ++r) acc += w[r];
  } else if (b < 0) {
    for (uint32_t r = 0; r < len; ++r) acc += w[r] * half_spin(lo[r], a);
  } else {
    for (uint32_t r = 0; r < len; ++r) acc += w[r] * (half_spin(lo[r], a) * half_spin(lo[r], b));
  }
  stats[(size_t)blockIdx.x * n_stats + q] = acc;
}

// out[0] = sum w; out[1 + i] = sum w s_i; out[1 + N + i*N + j] = sum w s_i s_j
// (sites i, j of the model; site i is bit N-1-i of the state).
__global__ __launch_bounds__(kBlock) void moment_combine_kernel(DevLayout L, const double* __restrict__ stats,
                                                                int n_stats, double* __restrict__ out) {
  const int N = L.N, B = L.B;
  const int e = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (e >= 1 + N + N * N) return;
  double acc = 0.0;
  if (e == 0) {
    for (uint32_t s = 0; s < L.n_seg; ++s) acc += stats[(size_t)s * n_stats];
  } else if (e <= N) {
    const int p = N - e;  // bit of site e-1
    for (uint32_t s = 0; s < L.n_seg; ++s) {
      const double* st = stats + (size_t)s * n_stats;
      acc += p >= B ? st[0] * half_spin(L.seg_hi[s], p - B) : st[1 + p];
    }
  } else {
    const int f = e - 1 - N;
    const int i = f / N, j = f % N;
    const int pi = N - 1 - i, pj = N - 1 - j;
    if (i == j) {
      for (uint32_t s = 0; s < L.n_seg; ++s) acc += 0.25 * stats[(size_t)s * n_stats];
    } else if (pi >= B && pj >= B) {
      for (uint32_t s = 0; s < L.n_seg; ++s) {
        const uint32_t h = L.seg_hi[s];
        acc += stats[(size_t)s * n_stats] * (half_spin(h, pi - B) * half_spin(h, pj - B));
      }
    } else if (pi >= B || pj >= B) {
      const int ph = pi >= B ? pi - B : pj - B;
      const int pl = pi >= B ? pj : pi;
      for (uint32_t s = 0; s < L.n_seg; ++s) acc += half_spin(L.seg_hi[s], ph) * stats[(size_t)s * n_stats + 1 + pl];
    } else {
      const int q = 1 + B + lo_pair_index(pi < pj ? pi : pj, pi < pj ? pj : pi, B);
      for (uint32_t s = 0; s < L.n_seg; ++s) acc += stats[(size_t)s * n_stats + q];
    }
  }
  out[e] = acc;
}

// Flip amplitudes of the pairs [blockIdx.y*8, +8): masks[2p] is the bit of
// site j (must be up), masks[2p+1] the bit of site i (must be down).
template <class V>
__global__ __launch_bounds__(kBlock) void flip_pairs_kernel(DevLayout L, const V* __restrict__ v,
                                                            const uint32_t* __restrict__ masks, int n_pairs,
                                                            double* __restrict__ part) {
  const int p0 = (int)blockIdx.y * kPairChunk;
  const int np = n_pairs - p0 < kPairChunk ? n_pairs - p0 : kPairChunk;
  uint32_t mu[kPairChunk], md[kPairChunk];
#pragma unroll
  for (int q = 0; q < kPairChunk; ++q) {
    mu[q] = q < np ? masks[2 * (p0 + q)] : 0u;
    md[q] = q < np ? masks[2 * (p0 + q) + 1] : 0u;
  }
  const uint32_t lomask = L.B >= 32 ? 0xffffffffu : ((1u << L.B) - 1u);
  double acc[kPairChunk];
#pragma unroll
  for (int q = 0; q < kPairChunk; ++q) acc[q] = 0.0;
  for (uint32_t s = blockIdx.x; s < L.n_seg; s += gridDim.x) {
    const uint32_t h = L.seg_hi[s];
    const int kl = L.k - __builtin_popcount(h);
    const uint32_t cls0 = L.lo_class_off[kl];
    const uint32_t len = L.lo_class_off[kl + 1] - cls0;
    const uint32_t base = L.hi_base[h];
    for (uint32_t r = threadIdx.x; r < len; r += kBlock) {
      const uint32_t st = (L.NH > 0 ? (h << L.B) : 0u) | L.lo_state[cls0 + r];
      const double x = (double)v[base + r];
#pragma unroll
      for (int q = 0; q < kPairChunk; ++q) {
        if (q < np && (st & mu[q]) != 0u && (st & md[q]) == 0u) {
          const uint32_t t = st ^ mu[q] ^ md[q];
          const uint32_t rr = L.hi_base[L.NH > 0 ? (t >> L.B) : 0u] + L.lo_rank[t & lomask];
          acc[q] += x * (double)v[rr];
        }
      }
    }
  }
#pragma unroll
  for (int q = 0; q < kPairChunk; ++q) {
    const double bs = block_sum_all(acc[q]);
    if (threadIdx.x == 0 && q < np) part[(size_t)(p0 + q) * gridDim.x + blockIdx.x] = bs;
  }
}

__global__ __launch_bounds__(kBlock) void flip_combine_kernel(const double* __restrict__ part, uint32_t nbx,
                                                              int n_pairs, double* __restrict__ out) {
  const int p = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (p >= n_pairs) return;
  double acc = 0.0;
  for (uint32_t b = 0; b < nbx; ++b) acc += part[(size_t)p * nbx + b];
  out[p] = acc;
}

}  // namespace

int moment_stat_count(const DevLayout& L) { return 1 + L.B + L.B * (L.B - 1) / 2; }

template <class V>
hipError_t launch_spin_moments(const DevLayout& L, const V* v, double* stats, double* out, hipStream_t st) {
  const int ns = moment_stat_count(L);
  if (ns > kBlock) return hipErrorInvalidValue;
  const size_t lds = (size_t)L.max_seg * (sizeof(double) + sizeof(uint32_t));
  hipLaunchKernelGGL((moment_stats_kernel<V>), dim3(L.n_seg), dim3(kBlock), lds, st, L, v, stats, ns);
  const int n_out = 1 + L.N + L.N * L.N;
  hipLaunchKernelGGL(moment_combine_kernel, dim3((n_out + kBlock - 1) / kBlock), dim3(kBlock), 0, st, L, stats, ns,
                     out);
  return hipGetLastError();
}

uint32_t flip_pair_blocks(const DevLayout& L) { return L.n_seg < 2048u ? L.n_seg : 2048u; }

template <class V>
hipError_t launch_flip_pairs(const DevLayout& L, const V* v, const uint32_t* masks, int n_pairs, double* part,
                             double* out, hipStream_t st) {
  if (n_pairs <= 0) return hipSuccess;
  const uint32_t nbx = flip_pair_blocks(L);
  const uint32_t nby = (uint32_t)((n_pairs + kPairChunk - 1) / kPairChunk);
  hipLaunchKernelGGL((flip_pairs_kernel<V>), dim3(nbx, nby), dim3(kBlock), 0, st, L, v, masks, n_pairs, part);
  hipLaunchKernelGGL(flip_combine_kernel, dim3((n_pairs + kBlock - 1) / kBlock), dim3(kBlock), 0, st, part, nbx,
                     n_pairs, out);
  return hipGetLastError();
}

template hipError_t launch_spin_moments<float>(const DevLayout&, const float*, double*, double*, hipStream_t);
template hipError_t launch_spin_moments<double>(const DevLayout&, const double*, double*, double*, hipStream_t);
template hipError_t launch_flip_pairs<float>(const DevLayout&, const float*, const uint32_t*, int, double*, double*,
                                             hipStream_t);
template hipError_t launch_flip_pairs<double>(const DevLayout&, const double*, const uint32_t*, int, double*,
                                              double*, hipStream_t);

}  // namespace alps_ed
