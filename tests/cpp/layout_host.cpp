// Host checks of the engine's layout and momentum-sector builders (no device):
//  - the diagonal tables (diag_lo, diag_hi and the cross bonds) against a
//    direct sum over sites and bonds for every state of random models;
//  - the lattice translations and the Bloch basis (representatives and
//    normalizations) against the oracle's restatement of
//    blochbasisstates.h for several lattices, sectors and momenta;
//  - the flip tables read with the H.v kernel's addressing (row record
//    (r/64)*64*n_pad + (r%64)*chunk, bond chunk k at k*64*chunk) against 4 *
//    the flipped words' ranks, for every class, row and bond of random
//    models, and every such read inside its class;
//  - the per-high-word flip lists against a recomputation from the states.
// Prints "diag_err <max abs error>", "bloch_sectors <count>" and
// "flip_bad <mismatches> flip_checked <entries>", or "bloch_mismatch ..."
// and exits 1.
#include <cmath>
#include <cstdio>
#include <random>

#include "../../oracle/ed_oracle.hpp"
#include "ed_bloch.hpp"
#include "ed_layout.hpp"

using namespace alps_ed;

static double diag_error() {
  std::mt19937 rng(5);
  std::uniform_real_distribution<double> U(-1, 1);
  double worst = 0;
  for (int trial = 0; trial < 48; ++trial) {
    const int N = 4 + trial % 18;
    ModelSpec s;
    s.n_sites = N;
    s.n_up = N / 2;
    const int nb = N + trial % 7;
    for (int b = 0; b < nb; ++b) {
      const int a = (int)(rng() % N), c = (int)((a + 1 + rng() % (N - 1)) % N);
      s.bonds.push_back(make_bond(a, c, U(rng), U(rng), U(rng)));
    }
    for (int i = 0; i < N; ++i) s.sites.push_back(make_site(U(rng), U(rng)));
    const SectorLayout L = build_layout(s);
    for (uint32_t st = 0; st < (1u << N); st += 1 + (uint32_t)(rng() % 5)) {
      double d = 0;
      for (int i = 0; i < N; ++i) d += ((st >> (N - 1 - i)) & 1u) ? s.sites[i].e_up : s.sites[i].e_dn;
      for (const BondTerm& b : s.bonds) {
        const uint32_t b1 = (st >> (N - 1 - b.s1)) & 1u, b2 = (st >> (N - 1 - b.s2)) & 1u;
        d += b1 == b2 ? b.e_par : b.e_anti;
      }
      const uint32_t lo = st & ((1u << L.B) - 1u), hi = L.NH ? st >> L.B : 0u;
      double got = L.diag_lo[lo] + L.diag_hi[hi];
      // diag_hi holds every cross bond's parallel value; the antiparallel
      // ones add anti - par, as the H.v kernel does
      for (const CrossBond& x : L.xb)
        if (((hi & x.hbit) != 0) != ((lo & x.lbit) != 0)) got += x.e_anti - x.e_par;
      worst = std::max(worst, std::fabs(got - d));
    }
  }
  return worst;
}

static int bloch_sectors() {
  struct Case {
    const char* name;
    int L, W;
  };
  const Case cases[] = {{"chain lattice", 10, 0}, {"square lattice", 4, 2}, {"ladder", 5, 0},
                        {"triangular lattice", 3, 3}, {"frustrated square lattice", 3, 3}};
  int checked = 0;
  for (const Case& c : cases) {
    Parameters p;
    p["L"] = std::to_string(c.L);
    if (c.W) p["W"] = std::to_string(c.W);
    const LatticeGraph g = named_lattice_graph(c.name, Evaluator(p));
    const edo::Graph og = edo::named_lattice(c.name, c.L, c.W, 0);
    const std::vector<std::vector<double>> ks = translation_momenta(g);
    if (ks != edo::translation_momenta(og)) {
      std::printf("bloch_mismatch momenta %s\n", c.name);
      return -1;
    }
    for (int n_up = 0; n_up <= g.n; ++n_up) {
      ModelSpec s;
      s.n_sites = g.n;
      s.n_up = n_up;
      for (const LatticeBond& b : g.bonds) s.bonds.push_back(make_bond(b.s, b.t, 1.0, 1.0, 0.0));
      for (int i = 0; i < g.n; ++i) s.sites.push_back(make_site(0.0, 0.0));
      const SectorLayout L = build_layout(s);
      BlochOrbits orb;
      for (size_t q = 0; q < ks.size(); ++q) {
        const Translations t = lattice_translations(g, ks[q]);
        const std::vector<edo::Translation> ot = edo::translations(og, ks[q]);
        if (t.perm.size() != ot.size()) {
          std::printf("bloch_mismatch translation count %s\n", c.name);
          return -1;
        }
        for (size_t i = 0; i < ot.size(); ++i)
          for (int site = 0; site < g.n; ++site)
            if ((size_t)t.perm[i][site] != ot[i].perm[site] || std::abs(t.phase[i] - ot[i].phase) > 1e-14) {
              std::printf("bloch_mismatch translation %s t=%zu\n", c.name, i);
              return -1;
            }
        if (q == 0) orb = bloch_orbits(L, t);
        const BlochSector b = bloch_sector(orb, t, g.n);
        const edo::BlochBasis ob = edo::build_bloch(g.n, n_up, ot);
        bool same = b.rep_state.size() == ob.reps.size();
        for (size_t i = 0; same && i < ob.reps.size(); ++i)
          same = b.rep_state[i] == edo::state_bits(ob.reps[i]) && std::fabs(b.norm[i] - ob.norm[i]) < 1e-14;
        if (!same) {
          std::printf("bloch_mismatch %s n_up=%d k=%zu: %zu vs %zu representatives\n", c.name, n_up, q,
                      b.rep_state.size(), ob.reps.size());
          return -1;
        }
        ++checked;
      }
    }
  }
  return checked;
}

static void flip_check(long& bad, long& checked) {
  std::mt19937 rng(11);
  for (int trial = 0; trial < 24; ++trial) {
    const int N = 6 + trial % 20;
    ModelSpec s;
    s.n_sites = N;
    s.n_up = N / 2;
    for (int b = 0; b < N + trial % 9; ++b) {
      const int a = (int)(rng() % N), c = (int)((a + 1 + rng() % (N - 1)) % N);
      s.bonds.push_back(make_bond(a, c, 0.5, 1.0, 0.0));
    }
    for (int i = 0; i < N; ++i) s.sites.push_back(make_site(0.0, 0.0));
    const SectorLayout L = build_layout(s);
    const int B = L.B;
    for (int c = 0; c <= B; ++c) {
      const uint32_t len = L.lo_class_off[c + 1] - L.lo_class_off[c];
      const size_t nll = L.nbp_off[c + 1] - L.nbp_off[c], nx = L.xn_off[c + 1] - L.xn_off[c];
      for (uint32_t r = 0; r < len; ++r) {
        const uint32_t lo = L.lo_state[L.lo_class_off[c] + r];
        const uint32_t lrec = (r >> 6) * 64u * L.n_ll_p + (r & 63u) * kLLChunk;
        const uint32_t xrec = (r >> 6) * 64u * L.n_xb_p + (r & 63u) * kXChunk;
        for (int b = 0; b < L.n_ll_p; ++b) {
          const size_t at = lrec + (size_t)(b / kLLChunk * kLLChunk) * 64u + b % kLLChunk;
          const uint32_t m = b < (int)L.ll.size() ? L.ll[b].mask : 0u, bits = lo & m;
          const uint32_t want = 8u * ((bits != 0u && bits != m) ? L.lo_rank[lo ^ m] : len);
          ++checked;
          if (at >= nll || L.nbp16[L.nbp_off[c] + at] != want) ++bad;
        }
        for (int b = 0; b < (int)L.xb.size(); ++b) {
          const size_t at = xrec + (size_t)(b / kXChunk * kXChunk) * 64u + b % kXChunk;
          const uint32_t lb = L.xb[b].lbit;
          const uint32_t want = 4u * L.lo_rank[lo ^ lb] | ((lo & lb) != 0u ? 0x8000u : 0u);
          ++checked;
          if (at >= nx || L.xn16[L.xn_off[c] + at] != want) ++bad;
        }
      }
    }
  }
}

// The H.v processing order is a permutation of the segments, and where the
// classes are dealt in runs, position p of a class reads run p % 8 of it in
// address order.
static long proc_check() {
  long bad = 0;
  const int cases[][2] = {{24, 12}, {30, 15}, {28, 13}, {26, 12}, {20, 10}, {27, 14}};
  for (const auto& c : cases) {
    ModelSpec s;
    s.n_sites = c[0];
    s.n_up = c[1];
    for (int i = 0; i < c[0]; ++i) {
      s.bonds.push_back(make_bond(i, (i + 1) % c[0], 1.0, 1.0, 0.0));
      s.sites.push_back(make_site(0.0, 0.0));
    }
    const SectorLayout L = build_layout(s);
    std::vector<uint32_t> a = L.seg_hi, b = L.proc_hi;
    std::sort(a.begin(), a.end());
    std::sort(b.begin(), b.end());
    if (a != b) ++bad;
    // within one class, the positions dealt to one XCD ascend in address
    std::vector<std::vector<uint32_t>> last(L.NH + 1, std::vector<uint32_t>(kXcds, 0));
    std::vector<std::vector<bool>> seen(L.NH + 1, std::vector<bool>(kXcds, false));
    if (L.seg_hi.size() >= 64u * kXcds && L.NH >= 3) {
      long order_bad = 0;
      for (size_t p = 0; p < L.proc_hi.size(); ++p) {
        const uint32_t h = L.proc_hi[p];
        const int pc = __builtin_popcount(h), g = (int)(p % kXcds);
        if (seen[pc][g] && h < last[pc][g]) ++order_bad;
        seen[pc][g] = true;
        last[pc][g] = h;
      }
      // the top-bit grouping (N=24) deals by hi >> (NH-3) instead: skip it
      if (c[0] != 24 && order_bad > (long)(L.NH + 1) * kXcds) ++bad;
    }
  }
  return bad;
}

// The per-high-word lists (host_lists): for every high word with a
// segment, its entries are the high-high bonds whose two sites differ in a
// state of that high word, in bond order, each with the first row of the
// state with both sites swapped (found by ranking that state's word, not
// from hi_base) and the bond's coefficient, then weight-0 entries of its own
// segment up to a multiple of four; and every cross bond's neighbour
// segment is the one of the state with the high site flipped, where such a
// state exists.  Random models; returns the mismatches.
static long list_check(long& checked) {
  std::mt19937 rng(11);
  std::uniform_real_distribution<double> U(-1, 1);
  long bad = 0;
  for (int trial = 0; trial < 24; ++trial) {
    const int N = 13 + trial % 10;
    ModelSpec s;
    s.n_sites = N;
    s.n_up = N / 2 - 1 + trial % 3;
    for (int b = 0; b < 2 * N; ++b) {
      const int a = (int)(rng() % N), c = (int)((a + 1 + rng() % (N - 1)) % N);
      s.bonds.push_back(make_bond(a, c, U(rng), U(rng), U(rng)));
    }
    for (int i = 0; i < N; ++i) s.sites.push_back(make_site(U(rng), U(rng)));
    const SectorLayout L = build_layout(s, -1, true, true);
    const uint32_t nhi = 1u << L.NH;
    // the first row of the segment of a high word: the rank of its smallest state
    auto first_row = [&](uint32_t h) -> uint32_t {
      const int kl = L.k - __builtin_popcount(h);
      if (kl < 0 || kl > L.B) return 0xffffffffu;
      const uint32_t lo = kl == 0 ? 0u : ((1u << kl) - 1u);
      const uint32_t st = (h << L.B) | lo;
      uint64_t r = 0;
      int m = 0;
      for (int p = 0; p < N; ++p)
        if (st >> p & 1u) r += binom(p, ++m);
      return (uint32_t)r;
    };
    for (uint32_t h = 0; h < nhi; ++h) {
      const uint32_t own = first_row(h);
      if (own != L.hi_base[h]) ++bad;
      if (own == 0xffffffffu) continue;
      std::vector<HHEntry> want;
      for (const FlipBond& f : L.hh) {
        const uint32_t bits = h & f.mask;
        if (bits != 0u && bits != f.mask) want.push_back({first_row(h ^ f.mask), 0u, f.t});
      }
      while (want.size() % 4 != 0) want.push_back({own, 0u, 0.0});
      const uint32_t e0 = L.hh_off[h], e1 = L.hh_off[h + 1];
      if (e1 - e0 != want.size()) {
        ++bad;
        continue;
      }
      for (size_t i = 0; i < want.size(); ++i, ++checked)
        if (L.hh_ent[e0 + i].nb != want[i].nb || L.hh_ent[e0 + i].t != want[i].t) ++bad;
      for (int c = 0; c < L.n_xb_p; ++c, ++checked) {
        const uint32_t hb = c < (int)L.xb.size() ? L.xb[c].hbit : 0u;
        const uint32_t nb = first_row(h ^ hb);
        if (L.xb_nb[(size_t)h * L.n_xb_p + c] != (nb == 0xffffffffu ? own : nb)) ++bad;
      }
    }
  }
  return bad;
}

// Parameter files and explicit graphs through the host parsers (the
// sanitizer build runs them too): tasks, evaluated couplings, graph bonds,
// and inputs that must be rejected with an error rather than misread.
static long parse_check() {
  long bad = 0;
  const std::string text =
      "LATTICE=\"chain lattice\"; L=8; J=1; Jz='J*0.5'; NUMBER_EIGENVALUES=2;\n"
      "{ Sz_total=0 }\n{ Sz_total=1; h=0.25 }\n";
  std::vector<Parameters> tasks = parse_input(text);
  const std::vector<Parameters> more = parse_input("GRAPH=\"4-site mixed\"; { J0=1; J1=0.5; h1=0.1 }\n");
  tasks.insert(tasks.end(), more.begin(), more.end());
  if (tasks.size() != 3) ++bad;
  for (const Parameters& p : tasks) {
    const TaskModel m = task_model(p);
    const TaskSectors ts = task_sectors(p, m.n_sites);
    if (m.s1.size() != m.jz.size() || ts.n_up.empty()) ++bad;
  }
  const std::string xml =
      "<?xml version=\"1.0\"?>\n<!-- graphs -->\n<LATTICES>\n"
      "<GRAPH name=\"g\" vertices=\"3\"><VERTEX id=\"2\" type=\"1\"><COORDINATE>0 1</COORDINATE></VERTEX>"
      "<EDGE source=\"1\" target=\"2\" type=\"3\"/><EDGE source=\"2\" target=\"3\"></EDGE></GRAPH>\n"
      "<GRAPH name='h'><EDGE source='1' target='5'/></GRAPH></LATTICES>\n";
  const auto gs = library_graphs(xml);
  if (gs.size() != 2 || gs[0].second.n != 3 || gs[0].second.bonds.size() != 2 || gs[0].second.site_type[1] != 1 ||
      gs[1].second.n != 5)
    ++bad;
  const char* rejected[] = {"<GRAPH name=\"x\"><EDGE source=\"1\"", "<GRAPH name=\"x\" vertices=\"2\"><VERTEX id=\"9\"/>",
                            "<GRAPH name=\"x\"><EDGE source=\"0\" target=\"1\"/></GRAPH>", "<GRAPH name=x>",
                            "<!-- open", "<GRAPH name=\"x\"><BOND/></GRAPH>"};
  for (const char* r : rejected) {
    try {
      library_graphs(r);
      ++bad;
    } catch (const std::invalid_argument&) {
    }
  }
  return bad;
}

int main() {
  std::printf("diag_err %.3e\n", diag_error());
  std::printf("parse_bad %ld\n", parse_check());
  std::printf("proc_bad %ld\n", proc_check());
  long list_checked = 0;
  const long list_bad = list_check(list_checked);
  std::printf("list_bad %ld list_checked %ld\n", list_bad, list_checked);
  long bad = 0, checked = 0;
  flip_check(bad, checked);
  std::printf("flip_bad %ld flip_checked %ld\n", bad, checked);
  const int n = bloch_sectors();
  if (n < 0) return 1;
  std::printf("bloch_sectors %d\n", n);
  return 0;
}
