// ed_device.hpp -- device-side building blocks shared by the single-solve
// and the batched kernels: wave/block reductions, uniform (scalar) table
// reads and the H.v of one row segment.
#pragma once

#include <hip/hip_runtime.h>

#include "ed_kernels.hpp"

namespace alps_ed {
namespace dev {

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Block sum, returned to every thread.
__device__ __forceinline__ double block_sum_all(double v) {
  __shared__ double red_smem[kBlock / 64];
  v = wave_sum(v);
  const int wave = threadIdx.x >> 6;
  __syncthreads();  // red_smem may still be read by a previous call
  if ((threadIdx.x & 63) == 0) red_smem[wave] = v;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int w = 0; w < kBlock / 64; ++w) s += red_smem[w];
  return s;
}

// Two block sums through one pair of barriers (each equal to its
// block_sum_all).
__device__ __forceinline__ void block_sum_all2(double& a, double& b) {
  __shared__ double red2_smem[2][kBlock / 64];
  a = wave_sum(a);
  b = wave_sum(b);
  const int wave = threadIdx.x >> 6;
  __syncthreads();  // red2_smem may still be read by a previous call
  if ((threadIdx.x & 63) == 0) {
    red2_smem[0][wave] = a;
    red2_smem[1][wave] = b;
  }
  __syncthreads();
  a = 0.0;
  b = 0.0;
#pragma unroll
  for (int w = 0; w < kBlock / 64; ++w) {
    a += red2_smem[0][w];
    b += red2_smem[1][w];
  }
}

// A thread's share of a sum of partials: entries tid, tid + kBlock, ... in
// index order.  The loads of eight terms are issued together before their
// additions (a loop with one load per addition waits for each in turn).
__device__ __forceinline__ double thread_partials(const Partials& P) {
  double v = 0.0;
  for (uint32_t i = threadIdx.x; i < P.n; i += 8u * kBlock) {
    double t[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) t[k] = i + k * kBlock < P.n ? P.p[i + k * kBlock] : 0.0;
#pragma unroll
    for (int k = 0; k < 8; ++k)
      if (i + k * kBlock < P.n) v += t[k];
  }
  return v;
}

__device__ __forceinline__ double sum_partials(const Partials& P) { return block_sum_all(thread_partials(P)); }

// The block's sum of v stored to *out by thread 0, for a block that makes
// one such call (its LDS slots are written once): one barrier.
__device__ __forceinline__ void store_block_sum_once(double* out, double v) {
  __shared__ double once_smem[kBlock / 64];
  v = wave_sum(v);
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) once_smem[wave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) s += once_smem[w];
    *out = s;
  }
}

__device__ __forceinline__ void store_partial(double* out, double v) {
  const double s = block_sum_all(v);
  if (threadIdx.x == 0) out[blockIdx.x] = s;
}


// Basis state of row i (binary search over the segment starts) and the row
// of state s (D when s is not in the sector): the split ranking
// rank = hi_base[s >> B] + lo_rank[s & (2^B - 1)].
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

// Uniform reads of the layout tables.  The tables are never written while a
// kernel runs, so they are read through the constant address space, which
// lets the compiler use scalar loads for wave-uniform addresses.
template <class T>
__device__ __forceinline__ T ldc(const T* p) {
  return *(const __attribute__((address_space(4))) T*)(p);
}

// Start-vector element i of stream `seed`: uniform on [lo, hi) from the
// splitmix64 output of the counter seed + (i+1) * golden gamma.
__device__ __forceinline__ uint64_t splitmix64_mix(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

__device__ __forceinline__ double start_value(uint64_t seed, uint64_t i, double lo, double hi) {
  const uint64_t z = splitmix64_mix(seed + (i + 1) * 0x9e3779b97f4a7c15ULL);
  const double u = (double)(z >> 11) * (1.0 / 9007199254740992.0);
  return lo + (hi - lo) * u;
}

// Whole struct read through the constant address space (scalar loads when
// the address is wave-uniform); T is plain data of 8-byte multiple size.
template <class T>
__device__ __forceinline__ T ldc_struct(const T* p) {
  static_assert(sizeof(T) % 8 == 0, "ldc_struct needs an 8-byte multiple");
  T out;
  uint64_t* o = reinterpret_cast<uint64_t*>(&out);
  const __attribute__((address_space(4))) uint64_t* s = (const __attribute__((address_space(4))) uint64_t*)(p);
#pragma unroll
  for (int i = 0; i < (int)(sizeof(T) / 8); ++i) o[i] = s[i];
  return out;
}

// A load of T at byte offset `off` (per lane, 32 bits) from a wave-uniform
// global base: the load adds the offset to the base, held in scalar
// registers, in the address unit (global_load with a scalar base), instead
// of a 64-bit vector add per lane and load.  The empty asm keeps the
// compiler from folding the offset into a 64-bit per-lane address.
template <class T>
__device__ __forceinline__ T ld_lane(const void* base, uint32_t off) {
  typedef __attribute__((address_space(1))) const char gchar;
  typedef __attribute__((address_space(1))) const T gT;
  asm volatile("" : "+v"(off));
  return *reinterpret_cast<gT*>(reinterpret_cast<gchar*>(reinterpret_cast<uintptr_t>(base)) + off);
}

// The same for an offset that serves many loads: lane_off() makes it opaque
// once, and ld_at / st_at use it as it is (no register copy per access).
__device__ __forceinline__ uint32_t lane_off(uint32_t off) {
  asm volatile("" : "+v"(off));
  return off;
}

template <class T>
__device__ __forceinline__ T ld_at(const void* base, uint32_t off) {
  typedef __attribute__((address_space(1))) const char gchar;
  typedef __attribute__((address_space(1))) const T gT;
  return *reinterpret_cast<gT*>(reinterpret_cast<gchar*>(reinterpret_cast<uintptr_t>(base)) + off);
}

template <class T>
__device__ __forceinline__ void st_at(void* base, uint32_t off, T v) {
  typedef __attribute__((address_space(1))) char gchar;
  typedef __attribute__((address_space(1))) T gT;
  *reinterpret_cast<gT*>(reinterpret_cast<gchar*>(reinterpret_cast<uintptr_t>(base)) + off) = v;
}

// Every multiply-add of a row is an explicit fma, so that rows and dot
// partials do not depend on what the compiler contracts.
//
// The rows of hv_segment for NQ row slots per thread (rows tid,
// tid + kBlock, ...), NQ fixed at compile time so that the loads of all
// slots are issued together.  Every slot of the wave but the last is full
// (hv_segment gives a wave the slots up to its last row); in the last, rows
// past the segment read row 0 and are not stored.  Stages the thread's rows
// of x in LDS and passes the workgroup barrier.  The flip bonds of each
// class are processed in chunks of CLL / CHH / CX (the table padding is a
// multiple of each).  Returns the thread's share of x.Hx when DOT.
//
// Addressing: a slot's element of a segment of x (or y) is read at the
// lane offset of slot 0 from a base moved by q * kBlock elements, or, for
// the last slot, at its own clamped offset; its flip-table records at slot
// 0's lane offset from a base moved by q * kBlock records.  Rows of the last
// slot past the segment stay inside the class table (padded to groups of
// 64 rows), whose padding entries are in-range offsets.
template <class V, bool DOT, int NQ, int CLL, int CHH, int CX>
__device__ __forceinline__ double hv_rows(const DevLayout& L, uint32_t h, const V* __restrict__ x,
                                          V* __restrict__ y, double* xs, uint32_t len, uint32_t base, double dh,
                                          const uint16_t* nbp, const uint16_t* xnp, const double* dlp, uint32_t e0,
                                          uint32_t e1) {
  constexpr uint32_t VS = (uint32_t)sizeof(V);
  constexpr int QL = NQ - 1;  // the last slot
  const uint32_t tid = threadIdx.x;
  const V* xb = x + base;
  static_assert(CLL % kLLChunk == 0 && CX % kXChunk == 0, "schedule chunks are multiples of the table chunks");
  typedef uint16_t u16x4 __attribute__((ext_vector_type(4)));
  typedef uint16_t u16x2 __attribute__((ext_vector_type(2)));
  const uint32_t rl = tid + QL * kBlock;
  const bool okl = rl < len;
  const uint32_t vo0 = lane_off(VS * tid);               // slot 0 (and every full slot)
  const uint32_t vol = lane_off(okl ? VS * rl : 0u);     // the last slot, clamped
  // flip-table records (flip_index, ed_layout.hpp): bond chunk k of row r
  // starts at rec(r) + k * 64 * chunk
  const uint32_t lo_ll = lane_off(2u * ((tid >> 6) * 64u * (uint32_t)L.n_ll_p + (tid & 63u) * kLLChunk));
  const uint32_t lo_x = lane_off(2u * ((tid >> 6) * 64u * (uint32_t)L.n_xb_p + (tid & 63u) * kXChunk));
  const uint32_t ll_slot = (uint32_t)kBlock * (uint32_t)L.n_ll_p;  // u16 entries per slot
  const uint32_t x_slot = (uint32_t)kBlock * (uint32_t)L.n_xb_p;
  // element of slot q of the segment starting at p
  auto seg_ld = [&](const V* p, int q) -> V {
    return q < QL ? ld_at<V>(p + q * kBlock, vo0) : ld_at<V>(p, vol);
  };
  V xo[NQ];
  double d[NQ], off[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    xo[q] = seg_ld(xb, q);
    d[q] = dh + (q < QL ? ld_at<double>(dlp + q * kBlock, vo0 * (8u / VS)) : ld_at<double>(dlp, vol * (8u / VS)));
    off[q] = 0.0;
  }
  // The cross-bond table entries of the first chunk are loaded now, so that
  // they arrive while the other bonds are processed.
  const int n_xb = (L.n_xb + CX - 1) / CX * CX;
  constexpr int KX = CX / kXChunk;
  u16x2 xr[NQ][KX];
  if (n_xb > 0) {
#pragma unroll
    for (int q = 0; q < NQ; ++q)
#pragma unroll
      for (int k = 0; k < KX; ++k) xr[q][k] = ld_at<u16x2>(xnp + q * x_slot + k * kXChunk * 64u, lo_x);
  }
  // low-low flips: target rows from the class table, values from LDS.  The
  // table entries of chunk b + CLL are loaded while chunk b is processed,
  // into the other of two register sets (ta, tb).
  const int n_ll = (L.n_ll + CLL - 1) / CLL * CLL;
  constexpr int KL = CLL / kLLChunk;
  u16x4 ta[NQ][KL], tb[NQ][KL];
  auto ll_load = [&](u16x4(&t)[NQ][KL], int b) {
#pragma unroll
    for (int q = 0; q < NQ; ++q)
#pragma unroll
      for (int k = 0; k < KL; ++k)
        t[q][k] = ld_at<u16x4>(nbp + q * ll_slot + (uint32_t)(b + k * kLLChunk) * 64u, lo_ll);
  };
  auto ll_apply = [&](const u16x4(&t)[NQ][KL], int b) {
    double tc[CLL];
#pragma unroll
    for (int c = 0; c < CLL; ++c) tc[c] = ldc(L.ll_t + b + c);
#pragma unroll
    for (int q = 0; q < NQ; ++q)
#pragma unroll
      for (int c = 0; c < CLL; ++c) {
        const uint32_t j = t[q][c / kLLChunk][c % kLLChunk];
        off[q] = fma(tc[c], *reinterpret_cast<const double*>(reinterpret_cast<const char*>(xs) + j), off[q]);
      }
  };
  if (n_ll > 0) ll_load(ta, 0);
  // The segment's slice of x into LDS: each thread stores its own rows (the
  // rows of the segment are exactly the rows of the threads' slots), with
  // the table loads above already in flight; then the barrier that every
  // wave of the workgroup passes once (hv_segment).
#pragma unroll
  for (int q = 0; q < QL; ++q) xs[tid + q * kBlock] = (double)xo[q];
  if (okl) xs[rl] = (double)xo[QL];
  if (tid == 0) xs[len] = 0.0;
  __syncthreads();
  for (int b0 = 0; b0 < n_ll; b0 += 2 * CLL) {
    if (b0 + CLL < n_ll) ll_load(tb, b0 + CLL);
    ll_apply(ta, b0);
    if (b0 + CLL >= n_ll) break;
    if (b0 + 2 * CLL < n_ll) ll_load(ta, b0 + 2 * CLL);
    ll_apply(tb, b0 + CLL);
  }
  // high-high flips: a bond flips every row of the segment or none (its two
  // bits lie in h), with a contiguous target run.  The bonds that flip h are
  // listed per high word in bond order (hh_off / hh_ent), padded with
  // weight-0 reads of this segment; slots of a chunk past the list do the
  // same.  Bonds that do not flip would add exact zeros, so the sums are
  // those over all bonds.
  {
    for (uint32_t e = e0; e < e1; e += CHH) {
      // neighbour segment starts as element offsets (a select between
      // pointers would make the loads generic-address ones)
      uint32_t xt[CHH];
      double tt[CHH];
#pragma unroll
      for (int c = 0; c < CHH; ++c) {
        if (c < kHHList || e + c < e1) {
          const HHEntry en = ldc_struct(L.hh_ent + e + c);
          xt[c] = en.nb;
          tt[c] = en.t;
        } else {
          xt[c] = base;
          tt[c] = 0.0;
        }
      }
      V v[NQ][CHH];
#pragma unroll
      for (int q = 0; q < NQ; ++q)
#pragma unroll
        for (int c = 0; c < CHH; ++c) v[q][c] = seg_ld(x + xt[c], q);
#pragma unroll
      for (int q = 0; q < NQ; ++q)
#pragma unroll
        for (int c = 0; c < CHH; ++c) off[q] = fma(tt[c], (double)v[q][c], off[q]);
    }
  }
  // cross bonds: diagonal always, flip when the high and low bits differ
  // (bit 15 of the table entry is the row's low bit).  Where the
  // neighbouring high word has no segment no row can flip (its low word
  // would need a set bit more or fewer than it has), and xb_nb points to this
  // segment.  The diagonal already holds every cross bond's parallel value
  // (diag_hi); each antiparallel bond adds its anti - par difference.
  const uint32_t* xnb = L.xb_nb + h * (uint32_t)L.n_xb_p;
  for (int b0 = 0; b0 < n_xb; b0 += CX) {
    uint32_t xt[CX], hx[CX];
    double de[CX], tx[CX];
#pragma unroll
    for (int c = 0; c < CX; ++c) {
      const XBondConst k = ldc_struct(L.xb_c + b0 + c);
      hx[c] = (h & k.hbit) != 0u ? 0x8000u : 0u;
      xt[c] = ldc(xnb + b0 + c);
      de[c] = k.de;
      tx[c] = k.t;
    }
    uint32_t e[NQ][CX];
#pragma unroll
    for (int q = 0; q < NQ; ++q)
#pragma unroll
      for (int k = 0; k < KX; ++k)
#pragma unroll
        for (int c = 0; c < kXChunk; ++c) e[q][k * kXChunk + c] = xr[q][k][c];
    if (b0 + CX < n_xb) {
#pragma unroll
      for (int q = 0; q < NQ; ++q)
#pragma unroll
        for (int k = 0; k < KX; ++k)
          xr[q][k] = ld_at<u16x2>(xnp + q * x_slot + (uint32_t)(b0 + CX + k * kXChunk) * 64u, lo_x);
    }
    V v[NQ][CX];
    bool f[NQ][CX];
#pragma unroll
    for (int q = 0; q < NQ; ++q)
#pragma unroll
      for (int c = 0; c < CX; ++c) {
        f[q][c] = ((e[q][c] ^ hx[c]) & 0x8000u) != 0u;
        v[q][c] = ld_lane<V>(x + xt[c], f[q][c] ? (e[q][c] & 0x7fffu) * (VS / 4u) : 0u);
      }
#pragma unroll
    for (int q = 0; q < NQ; ++q)
#pragma unroll
      for (int c = 0; c < CX; ++c) {
        // the flag as 1.0 / 0.0 and the value selected in the storage type
        // (one 32-bit select each; the asm keeps the select before the
        // conversion, where selecting the FP64 value takes two)
        V w = f[q][c] ? v[q][c] : (V)0;
        asm volatile("" : "+v"(w));
        d[q] = fma(f[q][c] ? 1.0 : 0.0, de[c], d[q]);
        off[q] = fma(tx[c], (double)w, off[q]);
      }
  }
  double dotp = 0.0;
  V* yb = y + base;
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    if (q < QL || okl) {
      const double xi = xs[q < QL ? tid + q * kBlock : rl];
      const double yv = fma(d[q], xi, off[q]);
      if (q < QL) st_at<V>(yb + q * kBlock, vo0, (V)yv);
      else st_at<V>(yb, vol, (V)yv);
      if (DOT) dotp = fma(xi, yv, dotp);  // FP64 value before the storage rounding
    }
  }
  return dotp;
}

// H.v over the row segment of record S (high word S.h; one workgroup of
// kBlock threads; dynamic LDS of at least (L.max_seg + 1) doubles):
//   y[r] = sum_c H[r][c] x[c] for the segment's rows r;
// with DOT, *dot_slot receives the segment's sum of x_r (H x)_r (FP64, before
// y is rounded to the storage type).
//
// The segment's slice of x is staged in LDS in FP64.  Each thread owns up to kRows
// rows (r = tid + q*kBlock); the number of row slots that hold rows is
// uniform over each wave and selects a fixed-size instance of the bond
// loops, so a wave does no work for slots past the segment (the waves of a
// short segment's last slot skip it, those of a segment of at most 64 rows
// after the first do nothing).  CLL, CHH,
// CX: chunk sizes of the flip loops (the schedule, ed_kernels.hpp).
template <class V, bool DOT, int CLL = 4, int CHH = 4, int CX = 2>
__device__ __forceinline__ void hv_segment_at(const DevLayout& L, const SegRec& S, const V* __restrict__ x,
                                              V* __restrict__ y, double* __restrict__ dot_slot) {
  extern __shared__ __align__(16) unsigned char smem[];
  // the segment's slice of x in FP64 (converted once per row, not once per
  // low-low flip that reads it); entry len is the zero slot
  double* xs = reinterpret_cast<double*>(smem);
  const uint32_t tid = threadIdx.x;
  const uint32_t len = S.len;
  const uint16_t* nbp = L.nbp16 + S.nbp;
  const uint16_t* xnp = L.xn16 + S.xnp;
  const double* dlp = L.diag_lo_cls + S.cls0;
  // row slots of this wave: slot q holds rows wave*64 + q*kBlock + lane
  const uint32_t w0 = tid & ~63u;
  const uint32_t nq = __builtin_amdgcn_readfirstlane(len > w0 ? (len - w0 + kBlock - 1) / kBlock : 0u);
  static_assert(kRows == 4, "hv_segment dispatches one to four row slots");
  // Every wave passes one workgroup barrier between the LDS stores of the
  // x slice and the first LDS reads: inside hv_rows, or here for a wave
  // without rows (the branch is uniform over each wave).
  double dotp = 0.0;
  if (nq == 0)
    __syncthreads();
  else if (nq == 1)
    dotp = hv_rows<V, DOT, 1, CLL, CHH, CX>(L, S.h, x, y, xs, len, S.base, S.dh, nbp, xnp, dlp, S.e0, S.e1);
  else if (nq == 2)
    dotp = hv_rows<V, DOT, 2, CLL, CHH, CX>(L, S.h, x, y, xs, len, S.base, S.dh, nbp, xnp, dlp, S.e0, S.e1);
  else if (nq == 3)
    dotp = hv_rows<V, DOT, 3, CLL, CHH, CX>(L, S.h, x, y, xs, len, S.base, S.dh, nbp, xnp, dlp, S.e0, S.e1);
  else
    dotp = hv_rows<V, DOT, 4, CLL, CHH, CX>(L, S.h, x, y, xs, len, S.base, S.dh, nbp, xnp, dlp, S.e0, S.e1);
  if (DOT) store_block_sum_once(dot_slot, dotp);
}

// The record of the segment of high word h, looked up in the tables.
__device__ __forceinline__ SegRec seg_record(const DevLayout& L, uint32_t h) {
  const int kl = L.k - __builtin_popcount(h);
  SegRec S;
  S.h = h;
  S.base = ldc(L.hi_base + h);
  S.cls0 = ldc(L.lo_class_off + kl);
  S.len = ldc(L.lo_class_off + kl + 1) - S.cls0;
  S.nbp = ldc(L.nbp_off + kl);
  S.xnp = ldc(L.xn_off + kl);
  S.e0 = ldc(L.hh_off + h);
  S.e1 = ldc(L.hh_off + h + 1);
  S.dh = ldc(L.diag_hi + h);
  S.pad = 0.0;
  return S;
}

// H.v over the row segment of high word h (hv_segment_at with the record
// looked up).
template <class V, bool DOT, int CLL = 4, int CHH = 4, int CX = 2>
__device__ __forceinline__ void hv_segment(const DevLayout& L, uint32_t h, const V* __restrict__ x,
                                           V* __restrict__ y, double* __restrict__ dot_slot) {
  hv_segment_at<V, DOT, CLL, CHH, CX>(L, seg_record(L, h), x, y, dot_slot);
}

// Row i of H.v in a momentum sector (vectors of 2 * Bl.dk reals, the real
// and imaginary parts of element i at 2i and 2i + 1); rows i >= Bl.dk do
// nothing.
// The diagonal is the site terms plus each bond's parallel / antiparallel
// value; each antiparallel bond flips to s' = s_i ^ mask, whose rank in the
// fixed-Sz sector (split ranking) gives its representative j and
// translation t, and contributes t_b * phase_t * norm_j / norm_i times x_j
// (hamiltonian_matrix.hpp:308-345 with bloch_basis_states::index_and_phase,
// blochbasisstates.h:47-55).  Returns Re(conj(x_i) (Hx)_i) when DOT.
template <class V, bool DOT>
__device__ __forceinline__ double bloch_row(const DevLayout& L, const DevBloch& Bl, const V* __restrict__ x,
                                            V* __restrict__ y, uint32_t i) {
  const uint32_t dk = Bl.dk;
  if (i >= dk) return 0.0;
  const uint32_t lomask = (1u << L.B) - 1u;
  const uint32_t s = Bl.rep_state[i];
  double d = 0.0;
  for (int site = 0; site < L.N; ++site) d += ((s >> (L.N - 1 - site)) & 1u) ? L.site_up[site] : L.site_dn[site];
  const double xr = (double)x[2 * i], xi = (double)x[2 * i + 1];
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
    const uint64_t pk = Bl.full_pk[rk];
    const uint32_t j = (uint32_t)pk;
    if (j == 0xffffffffu) continue;  // an orbit not in this sector
    const uint32_t tr = (uint32_t)(pk >> 32) & 0xffu;
    const double f = tb * (Bl.norm_tab[(uint32_t)(pk >> 40)] / ni);
    const double cr = f * Bl.ph_re[tr], ci = f * Bl.ph_im[tr];
    const double yr = (double)x[2 * j], yi = (double)x[2 * j + 1];
    ar += cr * yr - ci * yi;
    ai += cr * yi + ci * yr;
  }
  const double outr = d * xr + ar, outi = d * xi + ai;
  y[2 * i] = (V)outr;
  y[2 * i + 1] = (V)outi;
  return DOT ? xr * outr + xi * outi : 0.0;
}

}  // namespace dev
}  // namespace alps_ed
