// ed_bloch.hip -- H.v in a momentum (Bloch) sector.
//
// The sector's Hamiltonian is complex Hermitian, H = A + iB.  Vectors are
// stored as reals, the real and imaginary part of each element side by side
// (the layout of a std::complex array), and the product is
//   Re y = A Re x - B Im x,  Im y = B Re x + A Im x,
// the real symmetric form of H.  For Hermitian H the Lanczos coefficients
// are real, so the real recurrence on realified vectors is the complex
// recurrence of ietl::lanczos on the sector (same alpha, beta, T matrix):
// every other kernel of the engine applies unchanged.
//
// A single momentum sector's operator lists the flips of its rows once on
// the host (bloch_flip_rows) and streams them (bloch_list_kernel); where the
// list does not fit its packing, and in the batch, one thread per
// representative row searches for each flipped state's representative
// (dev::bloch_row, ed_device.hpp, which the batched kernels share).
#include <hip/hip_runtime.h>

#include "ed_device.hpp"
#include "ed_kernels.hpp"

namespace alps_ed {

namespace {

using namespace dev;

template <class V, bool DOT>
__global__ __launch_bounds__(kBlock) void bloch_hv_kernel(DevLayout L, DevBloch Bl, const V* __restrict__ x,
                                                          V* __restrict__ y, double* __restrict__ dot_out) {
  const double part = bloch_row<V, DOT>(L, Bl, x, y, blockIdx.x * kBlock + threadIdx.x);
  if (DOT) store_partial(dot_out, part);
}

// H.v over the listed flips (DevBloch::row_ptr/ent, bloch_flip_rows):
// kBlochLanes lanes per row, each adding every kBlochLanes-th entry of the
// row; the lanes' sums meet by shuffles and the first lane adds the
// diagonal and stores the row.  A wave's lanes read one contiguous run of
// entries (the rows of a wave are consecutive in the list); the gathers of
// x, one (re, im) pair per entry, are the only scattered reads.
template <class V, bool DOT>
__global__ __launch_bounds__(kBlock) void bloch_list_kernel(DevBloch Bl, const V* __restrict__ x,
                                                            V* __restrict__ y, double* __restrict__ dot_out) {
  typedef V V2 __attribute__((ext_vector_type(2)));
  typedef double D2 __attribute__((ext_vector_type(2)));
  constexpr uint32_t P = kBlochLanes;
  const V2* x2 = reinterpret_cast<const V2*>(x);
  const D2* cf = reinterpret_cast<const D2*>(Bl.coef);
  const uint32_t g = blockIdx.x * kBlock + threadIdx.x;
  const uint32_t i = g / P, sub = g % P;
  const uint32_t dk = Bl.dk;
  double ar = 0.0, ai = 0.0, br = 0.0, bi = 0.0;
  // the row's own terms are loaded first, next to the row bounds, so that
  // their latency overlaps the entry gathers
  double inv = 0.0, d = 0.0, xr = 0.0, xi = 0.0;
  if (i < dk) {
    inv = Bl.inv_norm[i];
    d = Bl.diag[i];
    const V2 xo = x2[i];
    xr = (double)xo.x;
    xi = (double)xo.y;
    const uint32_t k0 = Bl.row_ptr[i], k1 = Bl.row_ptr[i + 1];
    // two entries per lane and iteration, summed into two pairs
    uint32_t k = k0 + sub;
    for (; k + P < k1; k += 2 * P) {
      const uint32_t e0 = Bl.ent[k], e1 = Bl.ent[k + P];
      const D2 c0 = cf[e0 >> Bl.sh_c], c1 = cf[e1 >> Bl.sh_c];
      const V2 v0 = x2[e0 & Bl.mask_a], v1 = x2[e1 & Bl.mask_a];
      const double yr0 = (double)v0.x, yi0 = (double)v0.y, yr1 = (double)v1.x, yi1 = (double)v1.y;
      ar = fma(c0.x, yr0, fma(-c0.y, yi0, ar));
      ai = fma(c0.x, yi0, fma(c0.y, yr0, ai));
      br = fma(c1.x, yr1, fma(-c1.y, yi1, br));
      bi = fma(c1.x, yi1, fma(c1.y, yr1, bi));
    }
    if (k < k1) {
      const uint32_t e0 = Bl.ent[k];
      const D2 c0 = cf[e0 >> Bl.sh_c];
      const V2 v0 = x2[e0 & Bl.mask_a];
      const double yr0 = (double)v0.x, yi0 = (double)v0.y;
      ar = fma(c0.x, yr0, fma(-c0.y, yi0, ar));
      ai = fma(c0.x, yi0, fma(c0.y, yr0, ai));
    }
    ar += br;
    ai += bi;
  }
#pragma unroll
  for (uint32_t o = P / 2; o > 0; o >>= 1) {
    ar += __shfl_xor(ar, o, 64);
    ai += __shfl_xor(ai, o, 64);
  }
  double part = 0.0;
  if (i < dk && sub == 0) {
    const double outr = fma(d, xr, inv * ar), outi = fma(d, xi, inv * ai);
    V2 o;
    o.x = (V)outr;
    o.y = (V)outi;
    reinterpret_cast<V2*>(y)[i] = o;
    if (DOT) part = fma(xr, outr, xi * outi);
  }
  if (DOT) store_partial(dot_out, part);
}

// Configuration-basis form of a momentum-sector vector w (realified): on
// every state r of the fixed-Sz sector whose orbit is kept,
// phase_t(r) * norm_a * w_a with a its representative and t(r) the
// translation taking it there; 0 elsewhere.
template <class V>
__global__ __launch_bounds__(kBlock) void bloch_expand_kernel(DevBloch Bl, uint32_t D, const V* __restrict__ w,
                                                              double* __restrict__ re, double* __restrict__ im) {
  const uint32_t r = blockIdx.x * kBlock + threadIdx.x;
  if (r >= D) return;
  const uint64_t pk = Bl.full_pk[r];
  const uint32_t a = (uint32_t)pk;
  double vr = 0.0, vi = 0.0;
  if (a != 0xffffffffu) {
    const uint32_t t = (uint32_t)(pk >> 32) & 0xffu;
    const double pr = Bl.ph_re[t], pi = Bl.ph_im[t], nrm = Bl.norm[a];
    const double wr = (double)w[2 * a], wi = (double)w[2 * a + 1];
    vr = nrm * (pr * wr - pi * wi);
    vi = nrm * (pr * wi + pi * wr);
  }
  re[r] = vr;
  im[r] = vi;
}

}  // namespace

template <class V>
hipError_t launch_bloch_expand(const DevBloch& B, uint32_t D, const V* w, double* re, double* im, hipStream_t st) {
  if (D == 0) return hipSuccess;
  hipLaunchKernelGGL((bloch_expand_kernel<V>), dim3((D + kBlock - 1) / kBlock), dim3(kBlock), 0, st, B, D, w, re,
                     im);
  return hipGetLastError();
}

template hipError_t launch_bloch_expand<float>(const DevBloch&, uint32_t, const float*, double*, double*,
                                               hipStream_t);
template hipError_t launch_bloch_expand<double>(const DevBloch&, uint32_t, const double*, double*, double*,
                                                hipStream_t);

template <class V>
hipError_t launch_bloch_hv(const DevLayout& L, const DevBloch& B, const V* x, V* y, double* dot_out,
                           hipStream_t st) {
  if (B.dk == 0) return hipSuccess;
  const unsigned nb = bloch_blocks(B);
  if (B.row_ptr) {
    if (dot_out)
      hipLaunchKernelGGL((bloch_list_kernel<V, true>), dim3(nb), dim3(kBlock), 0, st, B, x, y, dot_out);
    else
      hipLaunchKernelGGL((bloch_list_kernel<V, false>), dim3(nb), dim3(kBlock), 0, st, B, x, y, dot_out);
  } else if (dot_out) {
    hipLaunchKernelGGL((bloch_hv_kernel<V, true>), dim3(nb), dim3(kBlock), 0, st, L, B, x, y, dot_out);
  } else {
    hipLaunchKernelGGL((bloch_hv_kernel<V, false>), dim3(nb), dim3(kBlock), 0, st, L, B, x, y, dot_out);
  }
  return hipGetLastError();
}

template hipError_t launch_bloch_hv<float>(const DevLayout&, const DevBloch&, const float*, float*, double*,
                                           hipStream_t);
template hipError_t launch_bloch_hv<double>(const DevLayout&, const DevBloch&, const double*, double*, double*,
                                            hipStream_t);

}  // namespace alps_ed
