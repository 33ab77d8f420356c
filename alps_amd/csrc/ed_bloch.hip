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
// One thread per representative row i (state s_i).  The diagonal is the
// site terms plus each bond's parallel / antiparallel value; each
// antiparallel bond flips to s' = s_i ^ mask, whose rank in the fixed-Sz
// sector (split ranking) gives its representative j and translation t, and
// contributes t_b * phase_t * norm_j / norm_i times x_j
// (hamiltonian_matrix.hpp:308-345 with bloch_basis_states::index_and_phase,
// blochbasisstates.h:47-55).  Gathers of x_j are scattered; the rows are
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
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  const uint32_t dk = Bl.dk;
  const uint32_t lomask = (1u << L.B) - 1u;
  double part = 0.0;
  if (i < dk) {
    const uint32_t s = Bl.rep_state[i];
    double d = 0.0;
    for (int site = 0; site < L.N; ++site) d += ((s >> (L.N - 1 - site)) & 1u) ? L.site_up[site] : L.site_dn[site];
    const double xr = (double)x[i], xi = (double)x[dk + i];
    const double ni = Bl.norm[i];
    double ar = 0.0, ai = 0.0;
    for (int b = 0; b < L.n_bonds; ++b) {
      const uint32_t p1 = L.bond_p1[b], p2 = L.bond_p2[b];
      if (((s >> p1) & 1u) == ((s >> p2) & 1u)) {
        d += L.bond_par[b];
        continue;
      }
      d += L.bond_anti[b];
      const double tb = L.bond_t[b];
      if (tb == 0.0) continue;
      const uint32_t t = s ^ ((1u << p1) | (1u << p2));
      const uint32_t rk = L.hi_base[L.NH > 0 ? (t >> L.B) : 0u] + L.lo_rank[t & lomask];
      const uint32_t j = Bl.full_rep[rk];
      if (j == 0xffffffffu) continue;  // an orbit not in this sector
      const uint32_t tr = Bl.full_tr[rk];
      const double f = tb * (Bl.norm[j] / ni);
      const double cr = f * Bl.ph_re[tr], ci = f * Bl.ph_im[tr];
      const double yr = (double)x[j], yi = (double)x[dk + j];
      ar += cr * yr - ci * yi;
      ai += cr * yi + ci * yr;
    }
    const double outr = d * xr + ar, outi = d * xi + ai;
    y[i] = (V)outr;
    y[dk + i] = (V)outi;
    if (DOT) part = xr * outr + xi * outi;
  }
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
