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

}  // namespace

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
