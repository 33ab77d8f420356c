// ed_batch.hpp -- many independent Lanczos solves advanced together.
//
// Each problem (one fixed-Sz sector of its own Hamiltonian, or one momentum
// sector of it) is a BatchJob in device memory.  One launch of each batched
// kernel serves every problem in the batch: its grid is a flat list of work
// items (problem, segment) for H.v -- (problem, block of kBlock
// representatives) in a momentum sector -- and (problem, row chunk) for the
// vector kernels, so a batch of small
// sectors fills the whole chip in one launch instead of leaving it idle
// between many tiny ones.  A Lanczos step of the batch is four launches
// (H.v with the alpha partials, alpha, update with the beta partials,
// beta), with no host round trip; the host only reads alpha and beta (from
// mapped host memory) when it tests convergence.
//
// The per-problem arithmetic (segments, chunks, partial sums and their
// order) depends only on the problem itself, never on the other problems
// in the batch: a problem's alpha, beta and eigenvalues are bit-identical
// whether it runs alone or in any mix.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "ed_kernels.hpp"

namespace alps_ed {

// One problem as the batched kernels see it.  The Lanczos vector u_j
// (v_j = scale[j] * u_j) lives in slot j % cap of a ring of cap slots of
// `stride` elements; with cap > jmax the whole basis is kept.
struct BatchJob {
  DevLayout L;
  DevBloch Bl;          // momentum sector (Bl.dk = 0: the whole fixed-Sz sector)
  uint32_t n;           // Lanczos vector length: L.D, or 2 * Bl.dk (realified)
  uint32_t n_hv;        // H.v work items and partials: segments or representative blocks
  void* vec;            // slot 0 (float or double, as the batch lane)
  uint64_t stride;      // elements between slots
  uint32_t cap;         // slots in the ring
  uint32_t n_chunks;    // row chunks of kBatchChunk rows
  int32_t j0;           // Lanczos index of step t = 0 of the launch group
  int32_t jmax;         // no step runs at j >= jmax
  uint64_t seed;        // start vector stream
  double* alpha;        // [jmax + 1]
  double* beta;         // [jmax + 1]
  double* scale;        // [jmax + 2]
  double* host_alpha;   // device view of mapped host mirrors
  double* host_beta;
  double* part_hv;      // [n_hv]      partials of the H.v dot
  double* part_up;      // [n_chunks]  partials of |r|^2
};

// Ritz vectors and their Rayleigh quotients for FP32 lanes:
//   out_i = sum_{j<m} coef[i*m + j] * scale[j] * u_j   (FP64, i < nv)
struct RitzJob {
  const void* vec;      // slot 0 of the kept basis
  uint64_t stride;
  uint32_t D;           // vector length
  uint32_t n_chunks;
  int32_t m, nv;
  const double* scale;
  const double* coef;   // [nv * m]
  double* out;          // [nv * D]
  double* part;         // [nv * n_chunks]  partials of |out_i|^2 (mapped host)
};

// y = H x over FP64 vectors, with per-item partials of x.Hx (items as the
// H.v of BatchJob).
struct PlainHvJob {
  DevLayout L;
  DevBloch Bl;
  const double* x;
  double* y;
  double* part;         // [n_hv] (mapped host)
};

struct BatchItem {
  uint32_t job;
  uint32_t idx;         // segment (processing order), representative block or row chunk
};

constexpr uint32_t kBatchChunk = (uint32_t)kBlock * 4;  // rows per vector work item

enum { kBatchStart = 0, kBatchAlpha = 1, kBatchBeta = 2 };

// Start vectors u_0 (uniform [-1,1) from the job's seed, the same values as
// launch_fill_random) with partials of |u_0|^2.
template <class V>
hipError_t launch_batch_start(const BatchJob* jobs, const BatchItem* items, uint32_t n_items, hipStream_t st);
// Step t of a launch group: w = H u_j into slot j+1, alpha partials.
template <class V>
hipError_t launch_batch_hv(const BatchJob* jobs, const BatchItem* items, uint32_t n_items, int t,
                           uint32_t max_seg, hipStream_t st);
// r = s_j w - alpha_j s_j u_j - beta_{j-1} s_{j-1} u_{j-1} in place of w.
template <class V>
hipError_t launch_batch_update(const BatchJob* jobs, const BatchItem* items, uint32_t n_items, int t,
                               hipStream_t st);
// One block per listed job: scale[0] (kBatchStart), alpha_j (kBatchAlpha)
// or beta_j and scale[j+1] (kBatchBeta), from the partials in index order.
hipError_t launch_batch_finish(const BatchJob* jobs, const uint32_t* job_list, uint32_t n_jobs, int t, int mode,
                               hipStream_t st);
template <class V>
hipError_t launch_batch_ritz(const RitzJob* jobs, const BatchItem* items, uint32_t n_items, hipStream_t st);
hipError_t launch_batch_plain_hv(const PlainHvJob* jobs, const BatchItem* items, uint32_t n_items,
                                 uint32_t max_seg, hipStream_t st);

}  // namespace alps_ed
