// ed_bloch.hip -- H.v in a momentum (Bloch) sector.
//
// The sector's Hamiltonian is complex Hermitian, H = A + iB.  Vectors are
// stored realified, (Re v, Im v), and the product is
//   Re y = A Re x - B Im x,  Im y = B Re x + A Im x,
// the real symmetric form of H.  For Hermitian H the Lanczos coefficients
// are real, so the real recurrence on realified vectors is the complex
// recurrence of ietl::lanczos on the sector (same alpha, beta, T matrix):
// every other kernel of the engine applies unchanged.
//
// One thread per representative row (dev::bloch_row, ed_device.hpp, which
// the batched kernels share).  Gathers of x_j are scattered; the rows are
// about D / N, so the kernel is small next to the sector's real H.v.
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
    const double wr = (double)w[a], wi = (double)w[Bl.dk + a];
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
  if (dot_out)
    hipLaunchKernelGGL((bloch_hv_kernel<V, true>), dim3(nb), dim3(kBlock), 0, st, L, B, x, y, dot_out);
  else
    hipLaunchKernelGGL((bloch_hv_kernel<V, false>), dim3(nb), dim3(kBlock), 0, st, L, B, x, y, dot_out);
  return hipGetLastError();
}

template hipError_t launch_bloch_hv<float>(const DevLayout&, const DevBloch&, const float*, float*, double*,
                                           hipStream_t);
template hipError_t launch_bloch_hv<double>(const DevLayout&, const DevBloch&, const double*, double*, double*,
                                            hipStream_t);

}  // namespace alps_ed
