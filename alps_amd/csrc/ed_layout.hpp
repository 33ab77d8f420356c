// ed_layout.hpp -- host-side description of one fixed-Sz sector of a spin-1/2
// Hamiltonian, in the form the gfx950 kernels consume.
//
// Basis order.  A state is a bit string of N bits, bit (N-1-s) set iff site s
// is up.  The reference orders the fixed-Sz basis lexicographically by the
// per-site index vector with site 0 most significant and index 0 = Sz=-1/2
// (src/alps/model/basisstates.h:176-278, sitebasisstates.h:75-104), which is
// ascending integer order of these bit strings among those with k set bits.
// The index of a state is therefore its colex rank
//     rank(s) = sum_{m=1..k} C(p_m, m),   p_1 < ... < p_k the set bits,
// the combinatorial number system.  It replaces the reference's binary search
// (basisstates.h:68-74) by table lookups.
//
// Split ranking.  The N bits are cut into a low part (bits 0..B-1) and a high
// part (bits B..N-1).  The first kl = popcount(lo) set bits are the low ones,
// so rank(s) = lo_rank[lo] + hi_base[hi] where
//     lo_rank[lo] = colex rank of lo among B-bit words with popcount kl,
//     hi_base[hi] = sum over set bits of hi at global position p, index m
//                   counted from kl+1, of C(p, m).
// All states sharing `hi` occupy one contiguous row segment of length
// C(B, k - popcount(hi)) starting at hi_base[hi]; within a segment the rows
// are the low words of that popcount in ascending order (lo_state).
//
// Matrix elements.  Spin-1/2 model of lib/xml/models.xml.in:342-373 with the
// local matrices of test/model/example3.output: per site -h*Sz + D*Sz^2, per
// bond Jz*Sz*Sz + Jxy*exchange_xy + K*biquadratic, i.e. a diagonal e_par
// (parallel spins), e_anti (antiparallel) and an off-diagonal t for the flip
// of an antiparallel pair.
#pragma once

#include <algorithm>
#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

namespace alps_ed {

constexpr int kMaxSites = 32;
constexpr int kXcds = 8;  // accelerator complex dies of an MI355X
constexpr int kClassSlices = 4;  // contiguous slices of a class swept in turn (uneven segment order)

// Binomial coefficients C(n, k) for 0 <= n <= 64 (Pascal's triangle, built
// once); 0 outside 0 <= k <= n.
inline uint64_t binom(int n, int k) {
  struct Table {
    uint64_t c[65][65];
    Table() {
      for (int i = 0; i <= 64; ++i)
        for (int j = 0; j <= 64; ++j) c[i][j] = j == 0 ? 1 : (i == 0 ? 0 : c[i - 1][j - 1] + c[i - 1][j]);
    }
  };
  static const Table t;
  if (k < 0 || n < 0 || k > n || n > 64) return 0;
  return t.c[n][k];
}

struct BondTerm {
  int s1, s2;           // sites, edge order of the graph
  double e_par;         // 0.25*Jz + 0.0625*K
  double e_anti;        // -0.25*Jz + 0.3125*K
  double t;             // 0.5*Jxy - 0.25*K
};

struct SiteTerm {
  double e_dn;          // -h*(-1/2) + D/4
  double e_up;          // -h*(+1/2) + D/4
};

struct ModelSpec {
  int n_sites = 0;
  int n_up = 0;
  std::vector<BondTerm> bonds;
  std::vector<SiteTerm> sites;
};

// Local matrix elements from couplings, written as the reference's two-term
// sums (each product is exact, and a two-term sum is order independent).
inline BondTerm make_bond(int s1, int s2, double jz, double jxy, double k) {
  BondTerm b;
  b.s1 = s1;
  b.s2 = s2;
  b.e_par = (0.25 * jz) + (0.0625 * k);
  b.e_anti = (-0.25 * jz) + (0.3125 * k);
  b.t = (0.5 * jxy) + (-0.25 * k);
  return b;
}

inline SiteTerm make_site(double h, double d) {
  SiteTerm s;
  s.e_dn = (-h * -0.5) + (d * 0.25);
  s.e_up = (-h * 0.5) + (d * 0.25);
  return s;
}

// Flip bond in bit coordinates.
struct FlipBond {
  uint32_t mask;   // both bits (positions in the word the kernel tests)
  uint32_t mask2;  // second word's bit for cross bonds (0 otherwise)
  double t;
};

// Cross bond: one bit in the high word, one in the low word.
struct CrossBond {
  uint32_t hbit;   // bit in the high word (hi-local coordinates)
  uint32_t lbit;   // bit in the low word
  double e_par, e_anti, t;
};

struct LowTables;
struct FlipTables;

// A listed high-high flip (the H.v kernel reads four with one scalar load).
struct HHEntry {
  uint32_t nb;   // first row of the neighbour segment
  uint32_t pad;
  double t;      // coefficient
};

// The per-bond constants of a cross bond as the H.v kernel reads them.
struct XBondConst {
  double de;      // e_anti - e_par
  double t;       // flip coefficient
  uint32_t hbit;  // bit in the high word (0 for a null bond)
  uint32_t pad;
};

// What the H.v kernel needs of the segment it processes, one record per
// processing index (built on the device for a single operator): its high
// word, first row, class offset and length, flip-table offsets, the range
// of its listed high-high flips and the high word's diagonal.  Read with
// one pair of scalar loads instead of two dependent rounds of table reads.
struct SegRec {
  uint32_t h, base, cls0, len, nbp, xnp, e0, e1;
  double dh, pad;
};

struct SectorLayout {
  int N = 0, k = 0, B = 0, NH = 0;
  uint64_t D = 0;
  // the shared tables the low-word and flip-table vectors below copy: the
  // same objects for every sector with the same low-word size and low-low /
  // cross-bond topology (lets a batch upload them once)
  const LowTables* low = nullptr;
  std::shared_ptr<const FlipTables> flip;
  std::vector<uint32_t> lo_rank;       // [2^B]
  std::vector<uint32_t> lo_state;      // [2^B] grouped by popcount, ascending
  std::vector<uint32_t> lo_class_off;  // [B+2]
  std::vector<uint32_t> hi_base;       // [2^NH]; 0xffffffff when no state has this hi
  std::vector<uint32_t> seg_hi;        // valid hi words, ascending
  std::vector<uint32_t> seg_start;     // first row of each segment
  std::vector<double> diag_lo_cls;     // diag_lo in lo_state order (coalesced)
  std::vector<uint32_t> proc_hi;       // hi words in H.v processing order (by class)
  uint32_t max_seg = 0;                // longest segment
  // H.v tables.  Bond lists are padded with null bonds (coefficient 0,
  // never active) to multiples of the kernel's chunk sizes.
  //   nbp16[nbp_off[c] + flip_index(r, b, n_ll_p, kLLChunk)]: 8 * the row
  //     reached from row r of a class-c segment by low-low bond b, or
  //     8 * C(B,c) (the zero slot) if not allowed: the byte offset of the
  //     row's FP64 element in the segment's LDS slice;
  //   xn16[xn_off[c] + flip_index(r, b, n_xb_p, kXChunk)]: 4 * the rank of
  //     lo ^ lbit_b in its class, with bit 15 = the low bit of cross bond b
  //     in row r: the byte offset of an FP32 element of x (twice it for FP64).
  // Both are used as byte offsets without a shift (C(12,6) * 8 < 2^16,
  // C(12,6) * 4 < 2^15).
  int n_ll_p = 0, n_hh_p = 0, n_xb_p = 0;
  std::vector<uint16_t> nbp16, xn16;
  std::vector<uint32_t> nbp_off, xn_off;  // [B+2]
  std::vector<double> ll_t;               // [n_ll_p]
  // The high-high bonds that flip the rows of high word h (the two bits of
  // h & mask differ), in bond order: entries [hh_off[h], hh_off[h+1]) of
  // hh_ent, each the first row of the segment of h ^ mask and the bond's
  // coefficient.  Each list is padded to a multiple of kHHList with
  // coefficient-0 entries that read h's own segment.
  std::vector<uint32_t> hh_off;           // [2^NH + 1] (empty when built on the device)
  std::vector<HHEntry> hh_ent;
  std::vector<uint32_t> hh_mask;          // [n_hh] per high-high bond: both bits
  std::vector<double> hh_t;               //   and its coefficient
  std::vector<uint32_t> xb_hbit;          // [n_xb_p], 0 = null
  std::vector<XBondConst> xb_c;           // [n_xb_p] per cross bond, null bonds zero
  // xb_nb[h * n_xb_p + c]: the first row of the segment of h ^ hbit_c, or of
  // h's own segment when that high word has none (no row of h flips then)
  std::vector<uint32_t> xb_nb;
  std::vector<double> diag_lo;         // [2^B] site + low-low bond diagonal
  std::vector<double> diag_hi;         // [2^NH] site + high-high bond diagonal + every cross bond's par value
  std::vector<FlipBond> ll, hh;        // low-low and high-high flip bonds
  std::vector<CrossBond> xb;           // cross bonds (all, for the diagonal)
  // reference-order description for the element enumeration
  std::vector<double> site_dn, site_up;          // per site
  std::vector<uint32_t> bond_p1, bond_p2;        // bit positions per bond
  std::vector<double> bond_par, bond_anti, bond_t;
};

// The H.v schedules (ed_kernels.hpp) process the flip bonds of each class in
// chunks of up to 8; bond lists and flip tables are padded to these
// multiples, which every schedule's chunk sizes divide.
constexpr int kPadLL = 8, kPadHH = 8, kPadX = 4;
// padding of each high word's list of flipping high-high bonds
constexpr int kHHList = 4;

inline int round_up(int n, int m) { return (n + m - 1) / m * m; }

// Flip tables hold each class in groups of 64 rows (one wave's rows).
// Within a group, the entries of bonds [k*ch, (k+1)*ch) of its 64 rows are
// contiguous, row after row: one load brings a row's ch entries, and the
// loads of a wave's 64 consecutive rows cover one contiguous run.  A class
// of len rows takes round_up(len, 64) * n_pad entries.
constexpr int kLLChunk = 4, kXChunk = 2;  // u16 entries per load (8 and 4 bytes)
static_assert(kPadLL % kLLChunk == 0 && kPadX % kXChunk == 0, "chunks divide the padding");

inline size_t flip_index(uint32_t r, int b, int n_pad, int ch) {
  return (size_t)(r >> 6) * 64u * n_pad + (size_t)(b / ch) * 64u * ch + (size_t)(r & 63u) * ch + b % ch;
}

inline int choose_low_bits(int N) {
  // A 12-bit low word gives row segments of up to C(12,6) = 924 rows: long
  // enough for coalesced streaming, short enough that a segment of x and the
  // 16-bit rank table (8 KiB) sit in LDS.  Smaller systems use all bits.
  return N < 12 ? N : 12;
}

// Tables of the B-bit low word, the same for every sector with that B:
// colex rank of each word within its popcount class, the words of each class
// in ascending order, and the class offsets.  Built once per B.
struct LowTables {
  std::vector<uint32_t> lo_rank, lo_state, lo_class_off;
};

inline const LowTables& low_tables(int B) {
  static std::mutex mu;
  static std::map<int, std::unique_ptr<LowTables>> cache;
  std::lock_guard<std::mutex> g(mu);
  std::unique_ptr<LowTables>& slot = cache[B];
  if (!slot) {
    slot.reset(new LowTables);
    LowTables& T = *slot;
    const uint32_t nlo = 1u << B;
    T.lo_rank.assign(nlo, 0);
    T.lo_class_off.assign(B + 2, 0);
    for (uint32_t v = 0; v < nlo; ++v) {
      uint64_t r = 0;
      int m_idx = 0;
      for (int p = 0; p < B; ++p)
        if (v >> p & 1u) r += binom(p, ++m_idx);
      T.lo_rank[v] = (uint32_t)r;
      T.lo_class_off[__builtin_popcount(v) + 1]++;
    }
    for (int c = 0; c <= B; ++c) T.lo_class_off[c + 1] += T.lo_class_off[c];
    T.lo_state.assign(nlo, 0);
    for (uint32_t v = 0; v < nlo; ++v) T.lo_state[T.lo_class_off[__builtin_popcount(v)] + T.lo_rank[v]] = v;
  }
  return *slot;
}

// Per-class flip tables of a set of low-word masks (low-low bonds, padded
// to n_pad with null bonds) and of cross-bond low bits, cached by their key:
// many sectors of one batch share the lattice and so these tables.
struct FlipTables {
  std::vector<uint32_t> nbp_off, xn_off;
  std::vector<uint16_t> nbp16, xn16;
};

inline std::shared_ptr<const FlipTables> flip_tables(int B, const std::vector<uint32_t>& ll_masks, int n_ll_p,
                                                     const std::vector<uint32_t>& x_lbits, int n_xb_p) {
  std::vector<uint32_t> key{(uint32_t)B, (uint32_t)n_ll_p, (uint32_t)n_xb_p, (uint32_t)ll_masks.size()};
  key.insert(key.end(), ll_masks.begin(), ll_masks.end());
  key.insert(key.end(), x_lbits.begin(), x_lbits.end());
  static std::mutex mu;
  static std::map<std::vector<uint32_t>, std::shared_ptr<const FlipTables>> cache;
  {
    std::lock_guard<std::mutex> g(mu);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
  }
  if (4 * binom(B, B / 2) >= 32768) throw std::invalid_argument("low word too wide for the flip tables");
  const LowTables& T = low_tables(B);
  std::shared_ptr<FlipTables> F(new FlipTables);
  F->nbp_off.assign(B + 2, 0);
  F->xn_off.assign(B + 2, 0);
  for (int c = 0; c <= B; ++c) {
    const uint64_t len64 = (uint64_t)round_up((int)binom(B, c), 64);
    F->nbp_off[c + 1] = F->nbp_off[c] + (uint32_t)(len64 * n_ll_p);
    F->xn_off[c + 1] = F->xn_off[c] + (uint32_t)(len64 * n_xb_p);
  }
  F->nbp16.assign(F->nbp_off[B + 1], 0);
  F->xn16.assign(F->xn_off[B + 1], 0);
  for (int c = 0; c <= B; ++c) {
    const uint32_t len = (uint32_t)binom(B, c);
    const uint32_t* ls = T.lo_state.data() + T.lo_class_off[c];
    uint16_t* nb = F->nbp16.data() + F->nbp_off[c];
    for (int b = 0; b < n_ll_p; ++b) {
      const uint32_t m = b < (int)ll_masks.size() ? ll_masks[b] : 0u;
      for (uint32_t r = 0; r < len; ++r) {
        const uint32_t lo = ls[r], bits = lo & m;
        nb[flip_index(r, b, n_ll_p, kLLChunk)] = (uint16_t)(8u * ((bits != 0u && bits != m) ? T.lo_rank[lo ^ m] : len));
      }
    }
    uint16_t* xn = F->xn16.data() + F->xn_off[c];
    for (int b = 0; b < (int)x_lbits.size(); ++b) {
      const uint32_t lb = x_lbits[b];
      for (uint32_t r = 0; r < len; ++r) {
        const uint32_t lo = ls[r];
        xn[flip_index(r, b, n_xb_p, kXChunk)] =
            (uint16_t)(4u * T.lo_rank[lo ^ lb] | (((lo & lb) != 0u ? 1u : 0u) << 15));
      }
    }
  }
  std::lock_guard<std::mutex> g(mu);
  if (cache.size() > 256) cache.clear();  // bounded: a batch uses a few lattices
  cache[key] = F;
  return F;
}

// copy_shared = false leaves lo_rank, lo_state, nbp16 and xn16 empty: the
// batch engine reads them from the shared tables (L.low, L.flip), which it
// uploads once per shape.  host_lists = false leaves the per-high-word lists
// (hh_off, hh_ent, xb_nb) empty: a single operator builds them on the device
// from hi_base and the bond masks (build_hv_lists), which avoids building
// and uploading up to 2^NH * (high-high + cross bonds) entries on the host.
inline SectorLayout build_layout(const ModelSpec& m, int B_override = -1, bool copy_shared = true,
                                 bool host_lists = false) {
  SectorLayout L;
  const int N = m.n_sites;
  if (N < 1 || N > kMaxSites) throw std::invalid_argument("number of sites must be in [1, 32]");
  if ((int)m.sites.size() != N) throw std::invalid_argument("site term count != number of sites");
  for (const BondTerm& b : m.bonds)
    if (b.s1 < 0 || b.s1 >= N || b.s2 < 0 || b.s2 >= N || b.s1 == b.s2)
      throw std::invalid_argument("bond with invalid sites");
  L.N = N;
  L.k = m.n_up;
  L.D = (m.n_up < 0 || m.n_up > N) ? 0 : binom(N, m.n_up);
  if (L.D >= 0xffffffffULL) throw std::invalid_argument("sector dimension exceeds 2^32-1");
  // segments longer than 1024 rows (kRows * kBlock of the H.v kernel) are
  // not supported, which bounds the low word to 12 bits
  const int B = B_override > 0 && B_override <= N && B_override <= 12 ? B_override : choose_low_bits(N);
  L.B = B;
  L.NH = N - B;
  const uint32_t nlo = 1u << B;
  const uint32_t nhi = 1u << L.NH;
  // low tables (depend on B only; shared)
  {
    const LowTables& T = low_tables(B);
    L.low = &T;
    if (copy_shared) {
      L.lo_rank = T.lo_rank;
      L.lo_state = T.lo_state;
    }
    L.lo_class_off = T.lo_class_off;
  }
  // high tables and segments (the set bits of each word in turn, binomials
  // from one row table: this loop runs 2^(N-B) times per layout)
  L.hi_base.assign(nhi, 0xffffffffu);
  {
    std::vector<std::vector<uint64_t>> cb(L.NH, std::vector<uint64_t>(N + 2, 0));
    for (int q = 0; q < L.NH; ++q)
      for (int m_idx = 0; m_idx <= N + 1; ++m_idx) cb[q][m_idx] = binom(B + q, m_idx);
    L.seg_hi.reserve(nhi);
    L.seg_start.reserve(nhi);
    for (uint32_t h = 0; h < nhi; ++h) {
      const int kl = L.k - __builtin_popcount(h);
      if (kl < 0 || kl > B || L.D == 0) continue;
      uint64_t r = 0;
      int m_idx = kl;
      for (uint32_t w = h; w != 0u; w &= w - 1u) r += cb[__builtin_ctz(w)][++m_idx];
      L.hi_base[h] = (uint32_t)r;
      L.seg_hi.push_back(h);
      L.seg_start.push_back((uint32_t)r);
    }
  }
  // segment processing order of the H.v kernel: longest segments first, so
  // that short ones fill the tail of the launch (popcount(hi) picks the
  // class; a longer class has kl nearer B/2), ascending hi within a length
  // (a stable counting sort by segment length)
  {
    std::vector<std::vector<uint32_t>> by_len(B + 1);  // by class kl, in ascending hi
    for (uint32_t h : L.seg_hi) by_len[L.k - __builtin_popcount(h)].push_back(h);
    std::vector<int> cls;
    for (int c = 0; c <= B; ++c)
      if (!by_len[c].empty()) cls.push_back(c);
    std::stable_sort(cls.begin(), cls.end(), [&](int a, int b) { return binom(B, a) > binom(B, b); });
    // classes of equal length interleave in ascending hi, as a stable sort
    // of the whole list by length leaves them
    L.proc_hi.clear();
    L.proc_hi.reserve(L.seg_hi.size());
    for (size_t i = 0; i < cls.size();) {
      size_t j = i;
      while (j < cls.size() && binom(B, cls[j]) == binom(B, cls[i])) ++j;
      // each class's list is ascending: merge them (at most two classes,
      // c and B - c, share a length)
      const size_t start = L.proc_hi.size();
      for (size_t q = i; q < j; ++q) {
        const size_t mid = L.proc_hi.size();
        L.proc_hi.insert(L.proc_hi.end(), by_len[cls[q]].begin(), by_len[cls[q]].end());
        std::inplace_merge(L.proc_hi.begin() + start, L.proc_hi.begin() + mid, L.proc_hi.end());
      }
      i = j;
    }
    if (!cls.empty()) L.max_seg = (uint32_t)binom(B, cls[0]);
  }
  // Workgroups are dealt to the 8 XCDs round-robin (block b runs on XCD
  // b % 8), each XCD with its own L2.  When the segments split evenly by the
  // top three bits of the high word, block b takes its segment from group
  // b % 8: most high-high neighbours (hi ^ mask with the top bits unchanged)
  // of a segment are then read on the XCD that owns them, from its L2.
  bool grouped = false;
  if (L.NH >= 3) {
    std::vector<std::vector<uint32_t>> grp(kXcds);
    for (uint32_t h : L.proc_hi) grp[h >> (L.NH - 3)].push_back(h);
    bool even = true;
    for (const auto& g : grp) even = even && g.size() == grp[0].size();
    if (even && !grp[0].empty()) {
      L.proc_hi.clear();
      for (size_t i = 0; i < grp[0].size(); ++i)
        for (int g = 0; g < kXcds; ++g) L.proc_hi.push_back(grp[g][i]);
      grouped = true;
    }
  }
  // Otherwise (popcount(hi) limited to a range narrower than 0..NH, so the
  // top bits split the segments unevenly) a large sector is dealt class by
  // class: each class, in address order, is cut into kXcds contiguous runs
  // of equal count, and the position p of the order takes from run p % 8,
  // i.e. XCD g sweeps run g of every class.  The high-high and cross
  // neighbours of a segment (hi ^ mask, near hi in address when the flipped
  // bits are low ones) are then read mostly on the XCD that reads hi.  Equal
  // counts of one class are equal work, so the XCDs stay balanced.  A class
  // is swept in kClassSlices contiguous slices one after the other, each
  // dealt as eight runs, which keeps the x of the segments in flight (and of
  // their neighbours) within a quarter of the class (J1-J2 6x5: 4.64 ->
  // 4.51 ms; 2 slices 4.58, 8 slices 4.50).
  if (!grouped && L.seg_hi.size() >= 64u * kXcds) {
    std::vector<int> cls_order;
    std::vector<std::vector<uint32_t>> by_pc(L.NH + 1);
    for (uint32_t h : L.proc_hi) {
      const int c = __builtin_popcount(h);
      if (by_pc[c].empty()) cls_order.push_back(c);
      by_pc[c].push_back(h);
    }
    L.proc_hi.clear();
    for (int c : cls_order) {
      std::vector<uint32_t>& v = by_pc[c];
      if (!std::is_sorted(v.begin(), v.end())) std::sort(v.begin(), v.end());  // address order: hi_base ascends with hi
      // kClassSlices contiguous slices in turn, each dealt as eight runs
      const size_t slice = (v.size() + kClassSlices - 1) / kClassSlices;
      for (size_t s0 = 0; s0 < v.size(); s0 += slice) {
        const size_t n = std::min(slice, v.size() - s0);
        const size_t base = n / kXcds, s = L.proc_hi.size();
        for (size_t i = 0; i < base; ++i)
          for (int j = 0; j < kXcds; ++j) L.proc_hi.push_back(v[s0 + ((s + j) % kXcds) * base + i]);
        for (size_t r = base * kXcds; r < n; ++r) L.proc_hi.push_back(v[s0 + r]);
      }
    }
  }
  // terms in bit coordinates
  auto pos = [N](int s) { return (uint32_t)(N - 1 - s); };
  // Diagonal tables by a recurrence over the highest set bit t of a word w:
  // diag[w] = diag[w ^ 2^t] + (a constant of bit t: its site term up - down
  // and its bonds to higher bits, all down) + for each bond (t, o) with o
  // below t the change from setting t (bit o read from w).  Every word costs
  // 1 + (bonds to lower bits) additions instead of one per site and bond.
  std::vector<double> su_lo(B, 0.0), su_hi(L.NH, 0.0);           // per bit constant
  std::vector<std::vector<std::pair<uint32_t, std::pair<double, double>>>> bl_lo(B), bl_hi(L.NH);
  double base_lo = 0.0, base_hi = 0.0;                              // every bit down
  for (int s = 0; s < N; ++s) {
    uint32_t p = pos(s);
    L.site_dn.push_back(m.sites[s].e_dn);
    L.site_up.push_back(m.sites[s].e_up);
    if ((int)p < B) {
      base_lo += m.sites[s].e_dn;
      su_lo[p] += m.sites[s].e_up - m.sites[s].e_dn;
    } else {
      base_hi += m.sites[s].e_dn;
      su_hi[p - B] += m.sites[s].e_up - m.sites[s].e_dn;
    }
  }
  for (const BondTerm& b : m.bonds) {
    uint32_t p1 = pos(b.s1), p2 = pos(b.s2);
    L.bond_p1.push_back(p1);
    L.bond_p2.push_back(p2);
    L.bond_par.push_back(b.e_par);
    L.bond_anti.push_back(b.e_anti);
    L.bond_t.push_back(b.t);
    bool lo1 = (int)p1 < B, lo2 = (int)p2 < B;
    // A bond's lower bit is set first in the recurrence (the higher one is
    // still down): par -> anti, a constant of that bit.  Setting the higher
    // bit then changes par -> anti when the lower bit is down and
    // anti -> par when it is up.
    if (lo1 && lo2) {
      base_lo += b.e_par;
      su_lo[std::min(p1, p2)] += b.e_anti - b.e_par;
      bl_lo[std::max(p1, p2)].push_back({1u << std::min(p1, p2), {b.e_anti - b.e_par, b.e_par - b.e_anti}});
      if (b.t != 0.0) L.ll.push_back({(1u << p1) | (1u << p2), 0u, b.t});
    } else if (!lo1 && !lo2) {
      uint32_t q1 = p1 - B, q2 = p2 - B;
      base_hi += b.e_par;
      su_hi[std::min(q1, q2)] += b.e_anti - b.e_par;
      bl_hi[std::max(q1, q2)].push_back({1u << std::min(q1, q2), {b.e_anti - b.e_par, b.e_par - b.e_anti}});
      if (b.t != 0.0) L.hh.push_back({(1u << q1) | (1u << q2), 0u, b.t});
    } else {
      uint32_t ph = lo1 ? p2 : p1, pl = lo1 ? p1 : p2;
      L.xb.push_back({1u << (ph - B), 1u << pl, b.e_par, b.e_anti, b.t});
    }
  }
  auto fill = [](std::vector<double>& d, uint32_t n, double base, const std::vector<double>& su,
                 const std::vector<std::vector<std::pair<uint32_t, std::pair<double, double>>>>& bl) {
    d.assign(n, 0.0);
    if (n == 0) return;
    d[0] = base;
    for (uint32_t w = 1; w < n; ++w) {
      const int t = 31 - __builtin_clz(w);
      const uint32_t u = w ^ (1u << t);
      double v = d[u] + su[t];
      for (const auto& e : bl[t]) v += (u & e.first) ? e.second.second : e.second.first;
      d[w] = v;
    }
  };
  fill(L.diag_lo, nlo, base_lo, su_lo, bl_lo);
  fill(L.diag_hi, nhi, base_hi, su_hi, bl_hi);
  // every cross bond's parallel value is part of every row's diagonal; the
  // kernel adds anti - par for the antiparallel ones
  double cross_par = 0.0;
  for (const CrossBond& c : L.xb) cross_par += c.e_par;
  for (double& v : L.diag_hi) v += cross_par;
  L.diag_lo_cls.resize(nlo);
  for (uint32_t i = 0; i < nlo; ++i) L.diag_lo_cls[i] = L.diag_lo[L.low->lo_state[i]];
  // padded H.v tables
  L.n_ll_p = round_up((int)L.ll.size(), kPadLL);
  L.n_hh_p = round_up((int)L.hh.size(), kPadHH);
  L.n_xb_p = round_up((int)L.xb.size(), kPadX);
  {
    std::vector<uint32_t> llm, xlb;
    for (const FlipBond& f : L.ll) llm.push_back(f.mask);
    for (const CrossBond& x : L.xb) xlb.push_back(x.lbit);
    std::shared_ptr<const FlipTables> F = flip_tables(B, llm, L.n_ll_p, xlb, L.n_xb_p);
    L.flip = F;
    L.nbp_off = F->nbp_off;
    L.xn_off = F->xn_off;
    if (copy_shared) {
      L.nbp16 = F->nbp16;
      L.xn16 = F->xn16;
    }
  }
  L.ll_t.assign(L.n_ll_p, 0.0);
  for (size_t b = 0; b < L.ll.size(); ++b) L.ll_t[b] = L.ll[b].t;
  L.hh_mask.assign(L.hh.size(), 0u);
  L.hh_t.assign(L.hh.size(), 0.0);
  for (size_t q = 0; q < L.hh.size(); ++q) {
    L.hh_mask[q] = L.hh[q].mask;
    L.hh_t[q] = L.hh[q].t;
  }
  // flipping high-high bonds per high word, in bond order: counted, then
  // written in place (branch-free: whether a bond flips is data-dependent,
  // and a mispredicted branch per bond cost more than the rest of the loop)
  if (host_lists) {
    const size_t nb = L.hh.size();
    std::vector<uint32_t> mk(nb);
    std::vector<double> tq(nb);
    for (size_t q = 0; q < nb; ++q) {
      mk[q] = L.hh[q].mask;
      tq[q] = L.hh[q].t;
    }
    auto flips = [&](uint32_t h, size_t q) -> uint32_t {
      const uint32_t bits = h & mk[q];
      return (uint32_t)(bits != 0u) & (uint32_t)(bits != mk[q]);
    };
    L.hh_off.assign((size_t)nhi + 1, 0u);
    uint64_t total = 0;
    for (uint32_t h = 0; h < nhi; ++h) {
      if (L.hi_base[h] != 0xffffffffu) {
        uint32_t n = 0;
        for (size_t q = 0; q < nb; ++q) n += flips(h, q);
        total += (uint64_t)round_up((int)n, kHHList);
      }
      L.hh_off[h + 1] = (uint32_t)std::min<uint64_t>(total, 0xffffffffULL);
    }
    if (total >= 0xffffffffULL) throw std::invalid_argument("high-high flip lists exceed 2^32-1 entries");
    // every bond's record is written at the list's next free entry, which
    // advances only when the bond flips; the entry past a full list is the
    // next high word's first (rewritten by it) or, after the last, one slack
    // entry dropped at the end
    L.hh_ent.resize(total + 1);
    for (uint32_t h = 0; h < nhi; ++h) {
      const uint32_t own = L.hi_base[h];
      if (own == 0xffffffffu) continue;
      HHEntry* e = L.hh_ent.data() + L.hh_off[h];
      HHEntry* const end = L.hh_ent.data() + L.hh_off[h + 1];
      for (size_t q = 0; q < nb; ++q) {
        *e = HHEntry{L.hi_base[h ^ mk[q]], 0u, tq[q]};
        e += flips(h, q);
      }
      for (; e < end; ++e) *e = HHEntry{own, 0u, 0.0};
    }
    L.hh_ent.pop_back();
  }
  L.xb_hbit.assign(L.n_xb_p, 0u);
  for (size_t b = 0; b < L.xb.size(); ++b) L.xb_hbit[b] = L.xb[b].hbit;
  L.xb_c.assign(L.n_xb_p, XBondConst{0.0, 0.0, 0u, 0u});
  for (size_t b = 0; b < L.xb.size(); ++b) L.xb_c[b] = {L.xb[b].e_anti - L.xb[b].e_par, L.xb[b].t, L.xb[b].hbit, 0u};
  if (host_lists) L.xb_nb.assign((size_t)nhi * L.n_xb_p, 0u);
  for (uint32_t h = 0; h < (host_lists ? nhi : 0u); ++h) {
    const uint32_t own = L.hi_base[h];
    if (own == 0xffffffffu) continue;
    for (int c = 0; c < L.n_xb_p; ++c) {
      const uint32_t nb = L.hi_base[h ^ L.xb_hbit[c]];
      L.xb_nb[(size_t)h * L.n_xb_p + c] = nb == 0xffffffffu ? own : nb;
    }
  }
  return L;
}

}  // namespace alps_ed
