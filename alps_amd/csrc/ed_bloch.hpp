// ed_bloch.hpp -- momentum (Bloch) sectors of the spin-1/2 path: lattice
// translations, the Bloch basis of a fixed-Sz sector and its device form.
//
// Mirrors:
//   - hypercubic_lattice::translations (src/alps/lattice/hypercubic.h:425-462)
//     and translation_momenta (:403-423): one translation per shift of the
//     periodic directions, the first direction fastest, with phase
//     exp(i sum_d shift_d k_d); momenta 2 pi j / extent per periodic
//     direction, the first direction outermost;
//   - bloch_basis_states::build (src/alps/model/blochbasisstates.h:97-271):
//     states in ascending order; a state's representative is the smallest
//     state of its orbit, reached by the first translation (in order) that
//     maps it there; a new representative is kept when
//     w = sqrt|1 + sum of the phases of the translations that fix it| >= 0.01,
//     with normalization w / sqrt(N); states of dropped orbits are not in
//     the basis;
//   - the matrix in that basis (hamiltonian_matrix.hpp:268-345): the element
//     from representative i to a state s' is
//     element * phase(s') * norm(rep(s')) / norm(i).
// The orbit structure (smallest state and the first translation reaching
// it, the translations fixing each representative) does not depend on the
// momentum; it is built once per sector (BlochOrbits) and each momentum
// selects its kept orbits (BlochSector).
#pragma once

#include <cmath>
#include <complex>
#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "ed_layout.hpp"
#include "ed_sparsediag.hpp"

namespace alps_ed {

struct Translations {
  std::vector<std::vector<int>> perm;  // perm[t][i]: image of site i
  std::vector<std::complex<double>> phase;
};

inline std::vector<int> periodic_directions(const LatticeGraph& g) {
  std::vector<int> dirs;
  for (size_t d = 0; d < g.extent.size(); ++d)
    if (g.periodic[d]) dirs.push_back((int)d);
  return dirs;
}

inline std::vector<std::vector<double>> translation_momenta(const LatticeGraph& g) {
  std::vector<std::vector<double>> ks;
  for (int d : periodic_directions(g)) {
    std::vector<std::vector<double>> next;
    if (ks.empty()) ks.push_back({});
    for (const auto& k : ks)
      for (int j = 0; j < g.extent[d]; ++j) {
        std::vector<double> v = k;
        v.push_back(2. * j * M_PI / g.extent[d]);
        next.push_back(v);
      }
    ks.swap(next);
  }
  return ks;
}

inline Translations lattice_translations(const LatticeGraph& g, const std::vector<double>& k) {
  const std::vector<int> dirs = periodic_directions(g);
  if (k.size() != dirs.size()) throw std::invalid_argument("the momentum needs one entry per periodic direction");
  const size_t dim = g.extent.size();
  Translations t;
  std::vector<int> shift(dirs.size(), 0);
  while (true) {
    double phase = 0.;
    std::vector<int> off(dim, 0);
    for (size_t i = 0; i < dirs.size(); ++i) {
      off[dirs[i]] = shift[i];
      phase += shift[i] * k[i];
    }
    std::vector<int> perm(g.n);
    for (int s = 0; s < g.n; ++s) {
      int idx = 0;
      for (size_t d = 0; d < dim; ++d) idx = idx * g.extent[d] + (g.offset[s][d] + off[d]) % g.extent[d];
      perm[s] = idx;
    }
    t.perm.push_back(perm);
    t.phase.push_back(std::exp(std::complex<double>(0., phase)));
    bool done = true;
    for (size_t i = 0; i < shift.size() && done; ++i) {
      if (++shift[i] >= g.extent[dirs[i]]) shift[i] = 0;
      else done = false;
    }
    if (done) break;
  }
  return t;
}

// Site permutation applied to bit-encoded states (site i is bit N-1-i)
// through one table per byte.
struct BitPermutation {
  std::vector<uint32_t> tab;  // [4][256]
  BitPermutation(int N, const std::vector<int>& perm) : tab(4 * 256, 0u) {
    for (int byte = 0; byte < 4; ++byte)
      for (uint32_t v = 0; v < 256; ++v) {
        uint32_t out = 0;
        for (int b = 0; b < 8; ++b) {
          const int bit = byte * 8 + b;
          if (((v >> b) & 1u) && bit < N) out |= 1u << (N - 1 - perm[N - 1 - bit]);
        }
        tab[byte * 256 + v] = out;
      }
  }
  uint32_t operator()(uint32_t s) const {
    return tab[s & 255u] | tab[256 + ((s >> 8) & 255u)] | tab[512 + ((s >> 16) & 255u)] | tab[768 + (s >> 24)];
  }
};

// Translations given as arrays (perm[t*N + i] = image of site i under
// translation t, phases re[t] + i im[t]); the first must be the identity.
inline Translations read_translations(int N, int n_t, const int* perm, const double* re, const double* im) {
  if (n_t < 1 || n_t > 64) throw std::invalid_argument("between 1 and 64 translations");
  if (!perm || !re || !im) throw std::invalid_argument("null argument");
  Translations t;
  for (int i = 0; i < n_t; ++i) {
    t.perm.emplace_back(perm + (size_t)i * N, perm + (size_t)(i + 1) * N);
    t.phase.emplace_back(re[i], im[i]);
  }
  for (int s = 0; s < N; ++s)
    if (t.perm[0][s] != s) throw std::invalid_argument("the first translation must be the identity");
  return t;
}

// Rank of state s (split ranking; the low-word ranks may live in the
// layout's shared tables).
inline uint32_t split_rank(const SectorLayout& L, uint32_t s) {
  const uint32_t h = L.NH > 0 ? (s >> L.B) : 0u;
  const std::vector<uint32_t>& lo = L.lo_rank.empty() ? L.low->lo_rank : L.lo_rank;
  return L.hi_base[h] + lo[s & ((1u << L.B) - 1u)];
}

// Momentum-independent orbit data of a fixed-Sz sector.
struct BlochOrbits {
  std::vector<uint32_t> state;     // per rank: the state (ascending)
  std::vector<uint32_t> rep_rank;  // per rank: rank of its orbit's smallest state
  std::vector<uint8_t> tr;         // per rank: first translation mapping it there (0 for the smallest)
  std::vector<uint32_t> reps;      // ranks of the smallest states, ascending
  std::vector<uint64_t> fixes;     // per entry of reps: translations t >= 1 that fix it (bit t)
};

inline BlochOrbits bloch_orbits(const SectorLayout& L, const Translations& t) {
  const int N = L.N;
  if (N > 32) throw std::invalid_argument("momentum sectors support at most 32 sites");
  if (t.perm.size() > 64) throw std::invalid_argument("at most 64 translations");
  for (size_t i = 0; i < t.perm.size(); ++i) {
    if ((int)t.perm[i].size() != N) throw std::invalid_argument("translation does not map every site");
    std::vector<int> seen(N, 0);
    for (int s : t.perm[i]) {
      if (s < 0 || s >= N || seen[s]) throw std::invalid_argument("translation is not a permutation of the sites");
      seen[s] = 1;
    }
  }
  std::vector<BitPermutation> bp;
  for (size_t i = 1; i < t.perm.size(); ++i) bp.emplace_back(N, t.perm[i]);
  const uint64_t D = L.D;
  BlochOrbits o;
  o.state.resize(D);
  o.rep_rank.resize(D);
  o.tr.resize(D);
  // fixed-Sz states in ascending order (the rank order)
  uint32_t s = L.k == 0 ? 0u : ((L.k >= 32 ? 0xffffffffu : ((1u << L.k) - 1u)));
  for (uint64_t r = 0; r < D; ++r) {
    o.state[r] = s;
    if (s != 0u) {  // next larger state with the same number of bits (Gosper)
      const uint32_t c = s & (0u - s), rr = s + c;
      s = (((rr ^ s) >> 2) / c) | rr;
    }
  }
  for (uint64_t r = 0; r < D; ++r) {
    const uint32_t x = o.state[r];
    if (split_rank(L, x) != r) throw std::logic_error("sector states are not in rank order");
    uint32_t best = x;
    int best_t = 0;
    uint64_t fix = 0;
    for (size_t i = 0; i < bp.size(); ++i) {
      const uint32_t y = bp[i](x);
      if (y == x) fix |= uint64_t(1) << (i + 1);
      else if (y < best) {  // the first translation reaching the smallest state
        best = y;
        best_t = (int)i + 1;
      }
    }
    if (best == x) {
      o.rep_rank[r] = (uint32_t)r;
      o.tr[r] = 0;
      o.reps.push_back((uint32_t)r);
      o.fixes.push_back(fix);
    } else {
      o.rep_rank[r] = split_rank(L, best);
      o.tr[r] = (uint8_t)best_t;
    }
  }
  return o;
}

constexpr uint32_t kNoState = 0xffffffffu;

// The Bloch basis at one momentum.  The device reads a flipped state's
// representative, translation and normalization with one gather of its
// packed entry: kept representative index (bits 0-31, kNoState when not
// kept), translation index (bits 32-39) and index into the table of
// distinct normalizations (bits 40-55).
struct BlochSector {
  std::vector<uint32_t> rep_state;  // kept representatives, ascending
  std::vector<double> norm;
  std::vector<uint32_t> full_rep;   // per rank: kept representative index, or kNoState
  std::vector<uint8_t> full_tr;     // per rank: translation index of its phase
  std::vector<uint64_t> full_pk;    // per rank: the packed entry
  std::vector<double> norm_tab;     // distinct normalizations
  std::vector<double> ph_re, ph_im; // per translation
};

inline BlochSector bloch_sector(const BlochOrbits& o, const Translations& t, int N) {
  BlochSector b;
  const double l = std::sqrt(double(N));
  std::vector<uint32_t> kept(o.state.size(), kNoState);  // by representative rank
  for (size_t q = 0; q < o.reps.size(); ++q) {
    std::complex<double> sum = 1.;
    for (size_t i = 1; i < t.phase.size(); ++i)
      if ((o.fixes[q] >> i) & 1u) sum += t.phase[i];
    const double w = std::sqrt(std::abs(sum));
    if (w < 0.01) continue;
    kept[o.reps[q]] = (uint32_t)b.rep_state.size();
    b.rep_state.push_back(o.state[o.reps[q]]);
    b.norm.push_back(w / l);
  }
  b.full_rep.resize(o.state.size());
  b.full_tr = o.tr;
  for (size_t r = 0; r < o.state.size(); ++r) b.full_rep[r] = kept[o.rep_rank[r]];
  std::map<double, uint32_t> tab;
  std::vector<uint32_t> norm_idx(b.norm.size());
  for (size_t q = 0; q < b.norm.size(); ++q) {
    auto it = tab.find(b.norm[q]);
    if (it == tab.end()) {
      if (b.norm_tab.size() >= 65535) throw std::invalid_argument("too many distinct Bloch normalizations");
      it = tab.emplace(b.norm[q], (uint32_t)b.norm_tab.size()).first;
      b.norm_tab.push_back(b.norm[q]);
    }
    norm_idx[q] = it->second;
  }
  b.full_pk.resize(o.state.size());
  for (size_t r = 0; r < o.state.size(); ++r) {
    const uint32_t a = b.full_rep[r];
    b.full_pk[r] = a == kNoState ? (uint64_t)kNoState
                                 : (uint64_t)a | ((uint64_t)b.full_tr[r] << 32) | ((uint64_t)norm_idx[a] << 40);
  }
  for (const auto& p : t.phase) {
    b.ph_re.push_back(p.real());
    b.ph_im.push_back(p.imag());
  }
  return b;
}

// The flips of every kept representative of a momentum sector, listed once
// so that the H.v streams them instead of searching for each flipped state's
// representative: for kept representative i (state s) in order, the bonds
// in edge order that are antiparallel in s and have a flip element, whose
// flipped state lies in a kept orbit a, give one entry: a and the code of
// the complex coefficient t_b * norm_a * phase_t (t the translation taking
// the flipped state to a); the row scales it by 1 / norm_i, which gives the
// element t_b * phase_t * norm_a / norm_i of hamiltonian_matrix.hpp:308-345
// with bloch_basis_states::index_and_phase (blochbasisstates.h:47-55).
// Coefficients are tabulated per (coupling, normalization, translation).
// The diagonal of row i is summed as the device row does it: the site terms
// in site order, then each bond's parallel or antiparallel value in bond
// order.  Entries are packed into 32 bits; `ok` is false when they do not
// fit (the device then searches per flip).
struct BlochFlipRows {
  bool ok = false;
  std::vector<uint32_t> row_ptr, ent;
  std::vector<double> diag, inv_norm, coef;  // coef: (re, im) per code
  uint32_t sh_c = 0, mask_a = 0;
};

inline int bits_for(uint64_t max_value) {
  int b = 1;
  while (b < 32 && (max_value >> b) != 0) ++b;
  return b;
}

inline BlochFlipRows bloch_flip_rows(const SectorLayout& L, const BlochSector& b) {
  BlochFlipRows r;
  const uint32_t dk = (uint32_t)b.rep_state.size();
  // coupling classes: the distinct nonzero flip elements
  std::map<double, uint32_t> cls;
  std::vector<uint32_t> bond_cls(L.bond_t.size(), 0);
  for (size_t q = 0; q < L.bond_t.size(); ++q)
    if (L.bond_t[q] != 0.0) {
      auto it = cls.emplace(L.bond_t[q], (uint32_t)cls.size()).first;
      bond_cls[q] = it->second;
    }
  std::vector<double> cls_t(cls.size(), 0.0);
  for (const auto& kv : cls) cls_t[kv.second] = kv.first;
  const uint64_t n_tr = b.ph_re.size(), n_ni = std::max<size_t>(b.norm_tab.size(), 1);
  const uint64_t n_code = std::max<uint64_t>(cls.size(), 1) * n_ni * n_tr;
  const int ba = bits_for(dk > 0 ? dk - 1 : 0), bc = bits_for(n_code - 1);
  if (ba + bc > 32 || n_code > (1u << 20)) return r;
  r.sh_c = (uint32_t)ba;
  r.mask_a = ba >= 32 ? 0xffffffffu : ((1u << ba) - 1u);
  r.coef.resize(2 * n_code);
  for (uint64_t c = 0; c < cls.size(); ++c)
    for (uint64_t n = 0; n < b.norm_tab.size(); ++n)
      for (uint64_t t = 0; t < n_tr; ++t) {
        const uint64_t code = (c * n_ni + n) * n_tr + t;
        const double f = cls_t[c] * b.norm_tab[n];
        r.coef[2 * code] = f * b.ph_re[t];
        r.coef[2 * code + 1] = f * b.ph_im[t];
      }
  r.row_ptr.reserve(dk + 1);
  r.row_ptr.push_back(0);
  r.diag.resize(dk);
  r.inv_norm.resize(dk);
  for (uint32_t i = 0; i < dk; ++i) {
    const uint32_t s = b.rep_state[i];
    double d = 0.0;
    for (int site = 0; site < L.N; ++site) d += ((s >> (L.N - 1 - site)) & 1u) ? L.site_up[site] : L.site_dn[site];
    for (size_t q = 0; q < L.bond_p1.size(); ++q) {
      const uint32_t p1 = L.bond_p1[q], p2 = L.bond_p2[q];
      if (((s >> p1) & 1u) == ((s >> p2) & 1u)) {
        d += L.bond_par[q];
        continue;
      }
      d += L.bond_anti[q];
      if (L.bond_t[q] == 0.0) continue;
      const uint32_t rk = split_rank(L, s ^ ((1u << p1) | (1u << p2)));
      const uint64_t pk = b.full_pk[rk];
      const uint32_t a = (uint32_t)pk;
      if (a == kNoState) continue;
      const uint64_t tr = (pk >> 32) & 0xffu, ni = (pk >> 40) & 0xffffu;
      const uint64_t code = ((uint64_t)bond_cls[q] * n_ni + ni) * n_tr + tr;
      r.ent.push_back(a | (uint32_t)(code << r.sh_c));
    }
    if (r.ent.size() > 0xffffffffull) return BlochFlipRows{};
    r.row_ptr.push_back((uint32_t)r.ent.size());
    r.diag[i] = d;
    r.inv_norm[i] = 1.0 / b.norm[i];
  }
  r.ok = true;
  return r;
}

// The orbits of a sector under a set of translations, built once and shared
// by every caller with the same sector shape and translations (the momenta
// of one Sz sector): one build even when several threads ask at once.
inline std::shared_ptr<const BlochOrbits> shared_bloch_orbits(const SectorLayout& L, const Translations& t) {
  struct Slot {
    std::once_flag once;
    std::shared_ptr<const BlochOrbits> orbits;
  };
  std::vector<int> key{L.N, L.k, L.B};
  for (const auto& p : t.perm) key.insert(key.end(), p.begin(), p.end());
  static std::mutex mu;
  static std::map<std::vector<int>, std::shared_ptr<Slot>> cache;
  static uint64_t cached_states = 0;
  std::shared_ptr<Slot> slot;
  {
    std::lock_guard<std::mutex> g(mu);
    auto it = cache.find(key);
    if (it != cache.end()) {
      slot = it->second;
    } else {
      // bounded: at most 64 sectors and 2^28 states (about 2.4 GB of orbit
      // data); a task has at most N + 1 Sz sectors
      if (cache.size() >= 64 || cached_states + L.D > (uint64_t(1) << 28)) {
        cache.clear();
        cached_states = 0;
      }
      slot = std::make_shared<Slot>();
      cache[key] = slot;
      cached_states += L.D;
    }
  }
  std::call_once(slot->once, [&] { slot->orbits = std::make_shared<const BlochOrbits>(bloch_orbits(L, t)); });
  return slot->orbits;
}

// The single matrix element of a momentum sector with one representative
// (the Bloch H.v row of ed_bloch.hip at row 0 applied to (1, 0)): the
// diagonal plus the flips that return to the representative's orbit, with
// their phases; the imaginary part cancels for a Hermitian H.
inline double bloch_single_element(const SectorLayout& L, const BlochSector& b) {
  const uint32_t s = b.rep_state.at(0);
  double d = 0.0, re = 0.0;
  for (int site = 0; site < L.N; ++site) d += ((s >> (L.N - 1 - site)) & 1u) ? L.site_up[site] : L.site_dn[site];
  for (size_t q = 0; q < L.bond_p1.size(); ++q) {
    const uint32_t p1 = L.bond_p1[q], p2 = L.bond_p2[q];
    if (((s >> p1) & 1u) == ((s >> p2) & 1u)) {
      d += L.bond_par[q];
      continue;
    }
    d += L.bond_anti[q];
    if (L.bond_t[q] == 0.0) continue;
    const uint32_t rk = split_rank(L, s ^ ((1u << p1) | (1u << p2)));
    if (b.full_rep[rk] != 0u) continue;  // another orbit, or one not in the sector
    re += L.bond_t[q] * b.ph_re[b.full_tr[rk]];
  }
  return d + re;
}

}  // namespace alps_ed
