// ed_batch_kernels.hip -- gfx950 kernels that advance many Lanczos solves
// in one launch (see ed_batch.hpp).  Every work item reads its problem's
// descriptor with wave-uniform (scalar) loads, then runs the same per-segment
// H.v and per-chunk vector code as the single-solve kernels.
#include <hip/hip_runtime.h>

#include "ed_batch.hpp"
#include "ed_device.hpp"

namespace alps_ed {

namespace {

using namespace dev;

template <class V>
__device__ __forceinline__ V* slot_ptr(void* vec, uint64_t stride, uint32_t cap, int j) {
  return static_cast<V*>(vec) + (uint64_t)((uint32_t)j % cap) * stride;
}

template <class V>
__global__ __launch_bounds__(kBlock) void batch_start_kernel(const BatchJob* __restrict__ jobs,
                                                             const BatchItem* __restrict__ items) {
  const BatchItem it = ldc_struct(items + blockIdx.x);
  const BatchJob* J = jobs + it.job;
  const uint32_t D = ldc(&J->n);
  const uint64_t seed = ldc(&J->seed);
  V* u = static_cast<V*>(ldc(&J->vec));
  double* part = ldc(&J->part_up);
  const uint32_t r0 = it.idx * kBatchChunk;
  const uint32_t r1 = min(D, r0 + kBatchChunk);
  double acc = 0.0;
  for (uint32_t i = r0 + threadIdx.x; i < r1; i += kBlock) {
    const V v = (V)start_value(seed, i, -1.0, 1.0);
    u[i] = v;
    acc += (double)v * (double)v;
  }
  const double s = block_sum_all(acc);
  if (threadIdx.x == 0) part[it.idx] = s;
}

template <class V>
__global__ __launch_bounds__(kBlock, HvWaves<V>::value) void batch_hv_kernel(const BatchJob* __restrict__ jobs,
                                                          const BatchItem* __restrict__ items, int t) {
  const BatchItem it = ldc_struct(items + blockIdx.x);
  const BatchJob* J = jobs + it.job;
  const int j = ldc(&J->j0) + t;
  if (j >= ldc(&J->jmax)) return;  // uniform over the workgroup
  const DevLayout L = ldc_struct(&J->L);
  void* vec = ldc(&J->vec);
  const uint64_t stride = ldc(&J->stride);
  const uint32_t cap = ldc(&J->cap);
  const V* x = slot_ptr<V>(vec, stride, cap, j);
  V* y = slot_ptr<V>(vec, stride, cap, j + 1);
  double* part = ldc(&J->part_hv);
  if (ldc(&J->Bl.dk) != 0u) {  // momentum sector: uniform over the workgroup
    const DevBloch Bl = ldc_struct(&J->Bl);
    const double p = bloch_row<V, true>(L, Bl, x, y, it.idx * kBlock + threadIdx.x);
    const double s = block_sum_all(p);
    if (threadIdx.x == 0) part[it.idx] = s;
    return;
  }
  const uint32_t h = ldc(L.proc_hi + it.idx);
  hv_segment<V, true>(L, h, x, y, part + it.idx);
}

template <class V>
__global__ __launch_bounds__(kBlock) void batch_update_kernel(const BatchJob* __restrict__ jobs,
                                                              const BatchItem* __restrict__ items, int t) {
  const BatchItem it = ldc_struct(items + blockIdx.x);
  const BatchJob* J = jobs + it.job;
  const int j = ldc(&J->j0) + t;
  if (j >= ldc(&J->jmax)) return;
  const uint32_t D = ldc(&J->n);
  void* vec = ldc(&J->vec);
  const uint64_t stride = ldc(&J->stride);
  const uint32_t cap = ldc(&J->cap);
  const V* u = slot_ptr<V>(vec, stride, cap, j);
  V* w = slot_ptr<V>(vec, stride, cap, j + 1);
  const double* alpha = ldc(&J->alpha);
  const double* beta = ldc(&J->beta);
  const double* scale = ldc(&J->scale);
  const double sc = ldc(scale + j);
  const double ca = ldc(alpha + j) * sc;
  const bool has_prev = j > 0;
  const V* up = has_prev ? slot_ptr<V>(vec, stride, cap, j - 1) : u;
  const double cb = has_prev ? ldc(beta + j - 1) * ldc(scale + j - 1) : 0.0;
  double* part = ldc(&J->part_up);
  // four consecutive rows per thread (slots are 256-byte aligned, chunks
  // start at multiples of kBatchChunk rows)
  const uint32_t i0 = it.idx * kBatchChunk + 4u * threadIdx.x;
  double acc = 0.0;
  if (i0 + 4u <= D) {
    typedef V v4 __attribute__((ext_vector_type(4), aligned(sizeof(V))));
    const v4 wv = *reinterpret_cast<const v4*>(w + i0);
    const v4 uv = *reinterpret_cast<const v4*>(u + i0);
    const v4 pv = has_prev ? *reinterpret_cast<const v4*>(up + i0) : uv;
    v4 o;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      double r = sc * (double)wv[c] - ca * (double)uv[c];
      if (has_prev) r -= cb * (double)pv[c];
      o[c] = (V)r;
      const double rd = (double)o[c];
      acc += rd * rd;
    }
    *reinterpret_cast<v4*>(w + i0) = o;
  } else {
    for (uint32_t i = i0; i < D && i < i0 + 4u; ++i) {
      double r = sc * (double)w[i] - ca * (double)u[i];
      if (has_prev) r -= cb * (double)up[i];
      const V rs = (V)r;
      w[i] = rs;
      const double rd = (double)rs;
      acc += rd * rd;
    }
  }
  const double s = block_sum_all(acc);
  if (threadIdx.x == 0) part[it.idx] = s;
}

__global__ __launch_bounds__(kBlock) void batch_finish_kernel(const BatchJob* __restrict__ jobs,
                                                              const uint32_t* __restrict__ job_list, int t,
                                                              int mode) {
  const BatchJob* J = jobs + ldc(job_list + blockIdx.x);
  const int j = ldc(&J->j0) + t;
  if (mode != kBatchStart && j >= ldc(&J->jmax)) return;
  const bool hv = mode == kBatchAlpha;
  const double* part = hv ? ldc(&J->part_hv) : ldc(&J->part_up);
  const uint32_t n = hv ? ldc(&J->n_hv) : ldc(&J->n_chunks);
  const double s = sum_partials(Partials{part, n});
  if (threadIdx.x != 0) return;
  double* scale = ldc(&J->scale);
  if (mode == kBatchStart) {
    scale[0] = 1.0 / sqrt(s);
  } else if (mode == kBatchAlpha) {
    const double sj = scale[j];
    const double a = sj * sj * s;
    ldc(&J->alpha)[j] = a;
    ldc(&J->host_alpha)[j] = a;
  } else {
    const double b = sqrt(s);
    ldc(&J->beta)[j] = b;
    ldc(&J->host_beta)[j] = b;
    scale[j + 1] = 1.0 / b;
  }
}

template <class V>
__global__ __launch_bounds__(kBlock) void batch_ritz_kernel(const RitzJob* __restrict__ jobs,
                                                            const BatchItem* __restrict__ items) {
  const BatchItem it = ldc_struct(items + blockIdx.x);
  const RitzJob* J = jobs + it.job;
  const V* vec = static_cast<const V*>(ldc(&J->vec));
  const uint64_t stride = ldc(&J->stride);
  const uint32_t D = ldc(&J->D);
  const int m = ldc(&J->m);
  const int nv = ldc(&J->nv);
  const double* scale = ldc(&J->scale);
  const double* coef = ldc(&J->coef);
  const uint32_t r0 = it.idx * kBatchChunk;
  const uint32_t r1 = min(D, r0 + kBatchChunk);
  double acc2[kMaxRitz];
#pragma unroll
  for (int i = 0; i < kMaxRitz; ++i) acc2[i] = 0.0;
  for (uint32_t r = r0 + threadIdx.x; r < r1; r += kBlock) {
    double acc[kMaxRitz];
#pragma unroll
    for (int i = 0; i < kMaxRitz; ++i) acc[i] = 0.0;
    for (int j = 0; j < m; ++j) {
      const double u = ldc(scale + j) * (double)vec[(uint64_t)j * stride + r];
#pragma unroll
      for (int i = 0; i < kMaxRitz; ++i)
        if (i < nv) acc[i] += ldc(coef + (size_t)i * m + j) * u;
    }
    double* out = ldc(&J->out);
#pragma unroll
    for (int i = 0; i < kMaxRitz; ++i)
      if (i < nv) {
        out[(uint64_t)i * D + r] = acc[i];
        acc2[i] += acc[i] * acc[i];
      }
  }
  double* part = ldc(&J->part);
  const uint32_t nc = ldc(&J->n_chunks);
#pragma unroll
  for (int i = 0; i < kMaxRitz; ++i) {
    if (i < nv) {  // uniform: every thread reaches the same barriers
      const double s = block_sum_all(acc2[i]);
      if (threadIdx.x == 0) part[(uint64_t)i * nc + it.idx] = s;
    }
  }
}

__global__ __launch_bounds__(kBlock) void batch_plain_hv_kernel(const PlainHvJob* __restrict__ jobs,
                                                                const BatchItem* __restrict__ items) {
  const BatchItem it = ldc_struct(items + blockIdx.x);
  const PlainHvJob* J = jobs + it.job;
  const DevLayout L = ldc_struct(&J->L);
  if (ldc(&J->Bl.dk) != 0u) {  // momentum sector: uniform over the workgroup
    const DevBloch Bl = ldc_struct(&J->Bl);
    const double p = bloch_row<double, true>(L, Bl, ldc(&J->x), ldc(&J->y), it.idx * kBlock + threadIdx.x);
    const double s = block_sum_all(p);
    if (threadIdx.x == 0) ldc(&J->part)[it.idx] = s;
    return;
  }
  const uint32_t h = ldc(L.proc_hi + it.idx);
  hv_segment<double, true>(L, h, ldc(&J->x), ldc(&J->y), ldc(&J->part) + it.idx);
}

}  // namespace

template <class V>
hipError_t launch_batch_start(const BatchJob* jobs, const BatchItem* items, uint32_t n_items, hipStream_t st) {
  if (n_items == 0) return hipSuccess;
  hipLaunchKernelGGL((batch_start_kernel<V>), dim3(n_items), dim3(kBlock), 0, st, jobs, items);
  return hipGetLastError();
}

template <class V>
hipError_t launch_batch_hv(const BatchJob* jobs, const BatchItem* items, uint32_t n_items, int t,
                           uint32_t max_seg, hipStream_t st) {
  if (n_items == 0) return hipSuccess;
  const size_t lds = (size_t)(max_seg + 1) * sizeof(double);  // the FP64 slice of hv_segment
  hipLaunchKernelGGL((batch_hv_kernel<V>), dim3(n_items), dim3(kBlock), lds, st, jobs, items, t);
  return hipGetLastError();
}

template <class V>
hipError_t launch_batch_update(const BatchJob* jobs, const BatchItem* items, uint32_t n_items, int t,
                               hipStream_t st) {
  if (n_items == 0) return hipSuccess;
  hipLaunchKernelGGL((batch_update_kernel<V>), dim3(n_items), dim3(kBlock), 0, st, jobs, items, t);
  return hipGetLastError();
}

hipError_t launch_batch_finish(const BatchJob* jobs, const uint32_t* job_list, uint32_t n_jobs, int t, int mode,
                               hipStream_t st) {
  if (n_jobs == 0) return hipSuccess;
  hipLaunchKernelGGL(batch_finish_kernel, dim3(n_jobs), dim3(kBlock), 0, st, jobs, job_list, t, mode);
  return hipGetLastError();
}

template <class V>
hipError_t launch_batch_ritz(const RitzJob* jobs, const BatchItem* items, uint32_t n_items, hipStream_t st) {
  if (n_items == 0) return hipSuccess;
  hipLaunchKernelGGL((batch_ritz_kernel<V>), dim3(n_items), dim3(kBlock), 0, st, jobs, items);
  return hipGetLastError();
}

hipError_t launch_batch_plain_hv(const PlainHvJob* jobs, const BatchItem* items, uint32_t n_items,
                                 uint32_t max_seg, hipStream_t st) {
  if (n_items == 0) return hipSuccess;
  const size_t lds = (size_t)(max_seg + 1) * sizeof(double);
  hipLaunchKernelGGL(batch_plain_hv_kernel, dim3(n_items), dim3(kBlock), lds, st, jobs, items);
  return hipGetLastError();
}

#define ALPS_ED_BATCH_INSTANTIATE(V)                                                                           \
  template hipError_t launch_batch_start<V>(const BatchJob*, const BatchItem*, uint32_t, hipStream_t);       \
  template hipError_t launch_batch_hv<V>(const BatchJob*, const BatchItem*, uint32_t, int, uint32_t,         \
                                         hipStream_t);                                                       \
  template hipError_t launch_batch_update<V>(const BatchJob*, const BatchItem*, uint32_t, int, hipStream_t); \
  template hipError_t launch_batch_ritz<V>(const RitzJob*, const BatchItem*, uint32_t, hipStream_t);

ALPS_ED_BATCH_INSTANTIATE(float)
ALPS_ED_BATCH_INSTANTIATE(double)

}  // namespace alps_ed
