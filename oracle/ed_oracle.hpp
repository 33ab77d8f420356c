// ed_oracle.hpp -- CPU FP64 restatement of the ALPS `sparsediag` hot path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the oracle: it is compiled into
// oracle/_build/liboracle_ed.so and oracle/_build/oracle_sparsediag, and only
// the tests, the smoke check and the CPU-baseline leg of the benchmark may load
// or run it.  The product (alps_amd/, include/) never links, includes or calls
// anything under oracle/.
//
// It restates, in plain single-threaded C++, the reference algorithm with the
// reference's own data structures where they matter for results:
//   * lattice graphs       src/alps/lattice/latticegraph.h:35-171,
//                          src/alps/lattice/hypercubic.h:110-200,253-280,
//                          lib/xml/lattices.xml.in:176-219,326-515
//   * spin model elements  lib/xml/models.xml.in:24-36,283-297,342-373,
//                          test/model/example3.output (HSpinSite, HSpinBond)
//   * basis states         src/alps/model/basisstates.h:176-278 (odometer,
//                          lexicographic order), :68-81 (lower_bound lookup)
//   * matrix build         src/alps/model/hamiltonian_matrix.hpp:44-58
//                          (explicit 0.0 diagonal), :269-345 (site/bond
//                          application), :388-397 (sites first, then bonds in
//                          edge order), src/alps/numeric/matrix_as_vector.hpp:73
//   * sparse H.v           src/ietl/interface/ublas.h:57-63 (row-wise, ascending
//                          column, map-of-maps storage as sparsediag.h:26)
//   * Lanczos              src/ietl/lanczos.h:341-406
//   * T-matrix analysis    src/ietl/tmatrix.h:100-228
//   * iteration control    src/ietl/iteration.h:67-134,209-216
//   * driver semantics     applications/diag/sparsediag/sparsediag.h:59-149
//
// Parity pins for every function are listed in tests/test_oracle.py.  The
// third-party arithmetic the reference delegates (LAPACK dstev, Boost.uBLAS
// norm_2 / inner_prod, Boost.Random lagged_fibonacci607) is restated here:
// dstev by an implicit-QL solver checked against scipy's LAPACK dstev
// fixtures; the start vector is a documented counter-based generator
// (splitmix64) instead of lagged_fibonacci607, so alpha/beta sequences are
// compared with the GPU path only, never with a reference run.
#pragma once

#include <algorithm>
#include <chrono>
#include <complex>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <limits>
#include <map>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace edo {

// ---------------------------------------------------------------------------
// Lattice graphs
// ---------------------------------------------------------------------------

struct UCEdge {
  int type;
  int sv;                 // source vertex in the unit cell
  std::vector<int> soff;  // source cell offset
  int tv;                 // target vertex in the unit cell
  std::vector<int> toff;  // target cell offset
};

struct UnitCell {
  int dim = 1;
  std::vector<int> vtype;  // one entry per unit-cell vertex
  std::vector<UCEdge> edges;
};

struct Bond {
  int s, t, type;
};

struct Graph {
  int n = 0;
  std::vector<int> stype;
  std::vector<Bond> bonds;
  std::vector<std::vector<int>> coord;  // cell offset of each site
  std::vector<int> ext;                 // extents of the finite lattice
  std::vector<bool> periodic;
  int cell_vertices = 1;
};

// Shift an offset by s on a finite lattice; false when it leaves an open
// boundary.  Periodic wrap in both directions (hypercubic.h:253-280).
inline bool lattice_shift(std::vector<int>& o, const std::vector<int>& s,
                          const std::vector<int>& ext, const std::vector<bool>& periodic) {
  for (size_t d = 0; d < ext.size(); ++d) {
    int v = o[d] + (d < s.size() ? s[d] : 0);
    if (v < 0) {
      if (!periodic[d]) return false;
      while (v < 0) v += ext[d];
    } else if (v >= ext[d]) {
      if (!periodic[d]) return false;
      v %= ext[d];
    }
    o[d] = v;
  }
  return true;
}

// Cell index: row-major, last dimension fastest (hypercubic.h:176-200).
inline int cell_index(const std::vector<int>& o, const std::vector<int>& ext) {
  int idx = 0;
  for (size_t d = 0; d < ext.size(); ++d) idx = idx * ext[d] + o[d];
  return idx;
}

// make_graph_from_lattice (latticegraph.h:35-171): cells in odometer order
// (last dimension fastest, hypercubic.h:118-140); per cell the unit-cell
// vertices, then the unit-cell edges in their listed order; an edge is kept
// when both ends lie on the lattice and are different sites.
inline Graph lattice_graph(const UnitCell& uc, const std::vector<int>& ext,
                           const std::vector<bool>& periodic) {
  Graph g;
  int volume = 1;
  for (int e : ext) volume *= e;
  const int nv = (int)uc.vtype.size();
  g.n = volume * nv;
  g.stype.resize(g.n);
  g.coord.resize(g.n);
  g.ext = ext;
  g.periodic = periodic;
  g.cell_vertices = nv;
  std::vector<int> cell(ext.size(), 0);
  for (int c = 0; c < volume; ++c) {
    for (int v = 0; v < nv; ++v) {
      g.stype[c * nv + v] = uc.vtype[v];
      g.coord[c * nv + v] = cell;
    }
    for (const UCEdge& e : uc.edges) {
      std::vector<int> os = cell, ot = cell;
      bool ins = lattice_shift(os, e.soff, ext, periodic);
      bool intg = lattice_shift(ot, e.toff, ext, periodic);
      if (!(ins && intg)) continue;
      int si = cell_index(os, ext) * nv + e.sv;
      int ti = cell_index(ot, ext) * nv + e.tv;
      if (si != ti) g.bonds.push_back({si, ti, e.type});
    }
    // advance the odometer, last dimension fastest
    for (int d = (int)ext.size() - 1; d >= 0; --d) {
      if (++cell[d] < ext[d]) break;
      cell[d] = 0;
    }
  }
  return g;
}

inline UCEdge uce(int type, std::vector<int> toff, size_t dim) {
  return UCEdge{type, 0, std::vector<int>(dim, 0), 0, std::move(toff)};
}

// The one-vertex unit cells and finite lattices of lib/xml/lattices.xml.in
// that the spin-1/2 path uses (:176-219 unit cells, :313-515 graphs).
inline Graph named_lattice(const std::string& name, int L, int W, int H) {
  UnitCell uc;
  uc.vtype = {0};
  std::vector<int> ext;
  std::vector<bool> per;
  // W defaults to L (2 for the ladders), H to W (lattices.xml.in:326-515)
  if (W <= 0) W = (name == "ladder" || name == "open ladder") ? 2 : L;
  if (H <= 0) H = W;
  if (name == "chain lattice" || name == "open chain lattice") {
    uc.dim = 1;
    uc.edges = {uce(0, {1}, 1)};
    ext = {L};
    per = {name == "chain lattice"};
  } else if (name == "nnn chain lattice" || name == "nnn open chain lattice") {
    uc.dim = 1;
    uc.edges = {uce(0, {1}, 1), uce(1, {2}, 1)};
    ext = {L};
    per = {name == "nnn chain lattice"};
  } else if (name == "square lattice" || name == "open square lattice") {
    uc.dim = 2;
    uc.edges = {uce(0, {0, 1}, 2), uce(0, {1, 0}, 2)};
    ext = {L, W};
    bool p = name == "square lattice";
    per = {p, p};
  } else if (name == "anisotropic square lattice") {
    uc.dim = 2;
    uc.edges = {uce(0, {1, 0}, 2), uce(1, {0, 1}, 2)};
    ext = {L, W};
    per = {true, true};
  } else if (name == "ladder" || name == "open ladder") {
    uc.dim = 2;
    uc.edges = {uce(0, {1, 0}, 2), uce(1, {0, 1}, 2)};
    ext = {L, W};
    per = {name == "ladder", false};
  } else if (name == "frustrated square lattice") {
    uc.dim = 2;
    uc.edges = {uce(0, {0, 1}, 2), uce(0, {1, 0}, 2), uce(1, {1, 1}, 2), uce(1, {1, -1}, 2)};
    ext = {L, W};
    per = {true, true};
  } else if (name == "triangular lattice") {
    uc.dim = 2;
    uc.edges = {uce(0, {0, 1}, 2), uce(0, {1, 0}, 2), uce(0, {-1, 1}, 2)};
    ext = {L, W};
    per = {true, true};
  } else if (name == "anisotropic triangular lattice") {
    uc.dim = 2;
    uc.edges = {uce(0, {0, 1}, 2), uce(1, {1, 0}, 2), uce(2, {-1, 1}, 2)};
    ext = {L, W};
    per = {true, true};
  } else if (name == "simple cubic lattice" || name == "open simple cubic lattice") {
    uc.dim = 3;
    uc.edges = {uce(0, {1, 0, 0}, 3), uce(0, {0, 1, 0}, 3), uce(0, {0, 0, 1}, 3)};
    ext = {L, W, H};
    bool p = name == "simple cubic lattice";
    per = {p, p, p};
  } else {
    throw std::runtime_error("oracle: unknown lattice '" + name + "'");
  }
  return lattice_graph(uc, ext, per);
}

// ---------------------------------------------------------------------------
// Spin-1/2 model: local matrices
// ---------------------------------------------------------------------------
// Site index 0 is Sz=-1/2, index 1 is Sz=+1/2 (sitebasisstates.h:75-104).

struct SpinModel {
  int n_sites = 0;
  std::vector<Bond> bonds;   // edge order
  std::vector<double> jz, jxy, k;  // per bond
  std::vector<double> h, d;        // per site
};

// Site term -h*Sz + D*Sz*Sz (models.xml.in:360-365), diagonal only for the
// Sz-conserving model; matches HSpinSite of test/model/example3.output.
inline double site_element(const SpinModel& m, int s, int is, int js) {
  if (is != js) return 0.0;
  double sz = is == 0 ? -0.5 : 0.5;
  return (-m.h[s] * sz) + (m.d[s] * (sz * sz));
}

// Bond term Jz*Sz*Sz + Jxy*exchange_xy + K*biquadratic, element
// <j1 j2|O|i1 i2> indexed mat[i1][i2][j1][j2]; matches HSpinBond of
// test/model/example3.output.
inline double bond_element(const SpinModel& m, int b, int i1, int i2, int j1, int j2) {
  double Jz = m.jz[b], Jxy = m.jxy[b], K = m.k[b];
  if (i1 == j1 && i2 == j2) {
    if (i1 == i2) return (0.25 * Jz) + (0.0625 * K);
    return (-0.25 * Jz) + (0.3125 * K);
  }
  if (i1 != i2 && j1 == i2 && j2 == i1) return (0.5 * Jxy) + (-0.25 * K);
  return 0.0;
}

// ---------------------------------------------------------------------------
// Basis states (basisstates.h:176-278)
// ---------------------------------------------------------------------------

typedef std::vector<short> State;

// Fixed-Sz basis in lexicographic order of the site-index vector (site 0 most
// significant, local index 0 before 1), i.e. the order the reference's
// odometer produces (last site fastest) with its constraint pruning
// (basisstates.h:223-260): a branch is cut as soon as the remaining sites can
// no longer reach the required number of up spins.
// n_up = number of sites in local state 1 (Sz_total = n_up - N/2).
inline void basis_dfs(int pos, int N, int need_up, State& cur, std::vector<State>& out) {
  if (pos == N) {
    if (need_up == 0) out.push_back(cur);
    return;
  }
  const int left_after = N - pos - 1;
  for (short v = 0; v < 2; ++v) {
    int nu = need_up - v;
    if (nu < 0 || nu > left_after) continue;
    cur[pos] = v;
    basis_dfs(pos + 1, N, nu, cur, out);
  }
}

inline std::vector<State> build_basis(int N, int n_up) {
  std::vector<State> states;
  if (N <= 0 || n_up < 0 || n_up > N) return states;
  State cur(N, 0);
  basis_dfs(0, N, n_up, cur, states);
  return states;
}

// std::lower_bound lookup; returns states.size() when absent (basisstates.h:68-74).
inline size_t basis_index(const std::vector<State>& states, const State& x) {
  auto it = std::lower_bound(states.begin(), states.end(), x);
  if (it == states.end() || !(*it == x)) return states.size();
  return (size_t)(it - states.begin());
}

// Bit encoding used to compare with the GPU path: bit (N-1-s) set iff site s
// is in local state 1 (spin up).  Ascending integer order equals the
// reference's lexicographic order within a fixed number of up spins.
inline uint32_t state_bits(const State& s) {
  uint32_t b = 0;
  int N = (int)s.size();
  for (int i = 0; i < N; ++i)
    if (s[i] == 1) b |= (1u << (N - 1 - i));
  return b;
}

// ---------------------------------------------------------------------------
// Hamiltonian build (hamiltonian_matrix.hpp:44-58, 269-345, 388-450)
// ---------------------------------------------------------------------------

// Row-major map-of-maps, as ublas::mapped_vector_of_mapped_vector (sparsediag.h:26).
typedef std::map<uint32_t, std::map<uint32_t, double>> MapMatrix;

// Build rows [row0, row1) of H (all rows when row1 == 0).
inline MapMatrix build_matrix(const SpinModel& m, const std::vector<State>& states,
                              size_t row0 = 0, size_t row1 = 0) {
  const size_t D = states.size();
  if (row1 == 0 || row1 > D) row1 = D;
  MapMatrix H;
  // initialized_matrix: explicit diagonal 0.0 via +1 then -1 (:44-58)
  for (size_t i = row0; i < row1; ++i) {
    H[(uint32_t)i][(uint32_t)i] += 1.;
    H[(uint32_t)i][(uint32_t)i] -= 1.;
  }
  // all site terms first, ascending site (:388-392, :269-291)
  for (int s = 0; s < m.n_sites; ++s) {
    for (size_t i = row0; i < row1; ++i) {
      const State& st = states[i];
      int is = st[s];
      for (int js = 0; js < 2; ++js) {
        double val = site_element(m, s, is, js);
        if (val != 0.0) {
          State ns = st;
          ns[s] = (short)js;
          size_t j = basis_index(states, ns);
          if (j < D) H[(uint32_t)i][(uint32_t)j] += val;
        }
      }
    }
  }
  // then all bonds in edge-index order (:394-396, :308-345)
  for (size_t b = 0; b < m.bonds.size(); ++b) {
    int s1 = m.bonds[b].s, s2 = m.bonds[b].t;
    for (size_t i = row0; i < row1; ++i) {
      const State& st = states[i];
      int is1 = st[s1], is2 = st[s2];
      for (int js1 = 0; js1 < 2; ++js1)
        for (int js2 = 0; js2 < 2; ++js2) {
          double val = bond_element(m, (int)b, is1, is2, js1, js2);
          if (val != 0.0) {
            State ns = st;
            ns[s1] = (short)js1;
            ns[s2] = (short)js2;
            size_t j = basis_index(states, ns);
            if (j < D) H[(uint32_t)i][(uint32_t)j] += val;
          }
        }
    }
  }
  return H;
}

// y = prod(H, x): per row, sum in ascending column order starting from 0.
inline void spmv(const MapMatrix& H, const double* x, double* y, size_t D) {
  for (size_t i = 0; i < D; ++i) y[i] = 0.0;
  for (const auto& row : H) {
    double t = 0.0;
    for (const auto& e : row.second) t += e.second * x[e.first];
    y[row.first] = t;
  }
}

// Flat CSR copy (row pointers, columns, values) of a map matrix.
struct CSR {
  std::vector<uint64_t> rowptr;
  std::vector<uint32_t> col;
  std::vector<double> val;
};

inline CSR to_csr(const MapMatrix& H, size_t row0, size_t row1) {
  CSR c;
  c.rowptr.push_back(0);
  for (size_t i = row0; i < row1; ++i) {
    auto it = H.find((uint32_t)i);
    if (it != H.end())
      for (const auto& e : it->second) {
        c.col.push_back(e.first);
        c.val.push_back(e.second);
      }
    c.rowptr.push_back(c.col.size());
  }
  return c;
}

inline void spmv_csr(const CSR& c, const double* x, double* y, size_t nrows) {
  for (size_t i = 0; i < nrows; ++i) {
    double t = 0.0;
    for (uint64_t p = c.rowptr[i]; p < c.rowptr[i + 1]; ++p) t += c.val[p] * x[c.col[p]];
    y[i] = t;
  }
}

// ---------------------------------------------------------------------------
// Eigenvector measurements (applications/diag/diag.h:191-314,
// sparsediag.h:151-159): <v| O |v> = inner_prod(v, prod(M, v)) with M the
// operator matrix of a site term A(i), a same-site product A(i)*B(i) or a
// two-site term A(i)*B(j), built as the Hamiltonian is (basis lookup of
// every target state, matrix_as_vector's M(row, target) += element,
// hamiltonian_matrix.hpp:269-345).
// ---------------------------------------------------------------------------
// Local 2x2 matrices m[out][in] = <out|A|in> of the spin-1/2 site operators
// of models.xml.in:24-36 (Splus, Sminus, Sz) and :283-285 (Sx =
// (Splus+Sminus)/2); index 0 is Sz=-1/2.
enum SiteOp { kOpSz = 0, kOpSplus = 1, kOpSminus = 2, kOpSx = 3 };

inline void site_op_matrix(int op, double m[2][2]) {
  m[0][0] = m[0][1] = m[1][0] = m[1][1] = 0.0;
  if (op == kOpSz) {
    m[0][0] = -0.5;
    m[1][1] = 0.5;
  } else if (op == kOpSplus) {
    m[1][0] = 1.0;  // sqrt(S(S+1) - Sz(Sz+1)) at Sz = -1/2
  } else if (op == kOpSminus) {
    m[0][1] = 1.0;
  } else if (op == kOpSx) {
    m[1][0] = 0.5;
    m[0][1] = 0.5;
  } else {
    throw std::invalid_argument("unknown site operator");
  }
}

// <v|O|v> for O = A(i) (j < 0), A(i)*B(i) (j == i, B applied first) or
// A(i)*B(j) (j != i).
inline double expect_site_term(const std::vector<State>& states, const double* v, int opA, int i, int opB,
                               int j) {
  const size_t D = states.size();
  double a[2][2], b[2][2];
  site_op_matrix(opA, a);
  if (j >= 0) site_op_matrix(opB, b);
  MapMatrix M;
  for (size_t r = 0; r < D; ++r) {
    const State& st = states[r];
    if (j < 0 || j == i) {
      double m[2][2];
      for (int o = 0; o < 2; ++o)
        for (int n = 0; n < 2; ++n) m[o][n] = j < 0 ? a[o][n] : a[o][0] * b[0][n] + a[o][1] * b[1][n];
      for (int js = 0; js < 2; ++js) {
        const double val = m[js][st[i]];
        if (val == 0.0) continue;
        State ns = st;
        ns[i] = (short)js;
        const size_t c = basis_index(states, ns);
        if (c < D) M[(uint32_t)r][(uint32_t)c] += val;
      }
    } else {
      for (int ji = 0; ji < 2; ++ji)
        for (int jj = 0; jj < 2; ++jj) {
          const double val = a[ji][st[i]] * b[jj][st[j]];
          if (val == 0.0) continue;
          State ns = st;
          ns[i] = (short)ji;
          ns[j] = (short)jj;
          const size_t c = basis_index(states, ns);
          if (c < D) M[(uint32_t)r][(uint32_t)c] += val;
        }
    }
  }
  std::vector<double> y(D);
  spmv(M, v, y.data(), D);
  double t = 0.0;
  for (size_t r = 0; r < D; ++r) t += v[r] * y[r];
  return t;
}

// ---------------------------------------------------------------------------
// Momentum sectors: lattice translations (hypercubic.h:425-462 over the
// periodic directions, latticegraph.h:252-268), the Bloch basis
// (blochbasisstates.h:97-271) and the Hamiltonian in it
// (hamiltonian_matrix.hpp:268-345: element * phase * norm(j) / norm(i)).
// ---------------------------------------------------------------------------
struct Translation {
  std::complex<double> phase;
  std::vector<size_t> perm;  // site i goes to perm[i]
};

// One translation per shift of the periodic directions (first direction
// fastest), phase exp(i sum shift_d k_d); k has one entry per periodic
// direction.
inline std::vector<Translation> translations(const Graph& g, const std::vector<double>& k) {
  std::vector<int> dirs;
  for (size_t d = 0; d < g.ext.size(); ++d)
    if (g.periodic[d]) dirs.push_back((int)d);
  if (k.size() != dirs.size()) throw std::runtime_error("Incorrect number of momenta specified");
  std::vector<Translation> out;
  std::vector<int> shift(dirs.size(), 0);
  const int nv = g.cell_vertices;
  while (true) {
    double phase = 0.;
    std::vector<int> off(g.ext.size(), 0);
    for (size_t i = 0; i < dirs.size(); ++i) {
      off[dirs[i]] = shift[i];
      phase += shift[i] * k[i];
    }
    Translation t;
    t.phase = std::exp(std::complex<double>(0., phase));
    for (int c = 0; c < g.n / nv; ++c) {
      std::vector<int> o = g.coord[c * nv];
      if (!lattice_shift(o, off, g.ext, g.periodic)) throw std::logic_error("shift leaves the lattice");
      for (int v = 0; v < nv; ++v) t.perm.push_back((size_t)(cell_index(o, g.ext) * nv + v));
    }
    out.push_back(t);
    bool done = true;
    for (size_t i = 0; i < shift.size() && done; ++i) {
      if (++shift[i] >= g.ext[dirs[i]]) shift[i] = 0;
      else done = false;
    }
    if (done) break;
  }
  return out;
}

// translation_momenta (hypercubic.h:403-423): 2 pi j / extent per periodic
// direction, the first direction outermost.
inline std::vector<std::vector<double>> translation_momenta(const Graph& g) {
  std::vector<std::vector<double>> ks;
  for (size_t d = 0; d < g.ext.size(); ++d) {
    if (!g.periodic[d]) continue;
    if (ks.empty()) {
      for (int j = 0; j < g.ext[d]; ++j) ks.push_back({2. * j * M_PI / g.ext[d]});
    } else {
      std::vector<std::vector<double>> nk;
      for (const auto& k : ks)
        for (int j = 0; j < g.ext[d]; ++j) {
          std::vector<double> v = k;
          v.push_back(2. * j * M_PI / g.ext[d]);
          nk.push_back(v);
        }
      ks.swap(nk);
    }
  }
  return ks;
}

struct BlochBasis {
  std::vector<State> reps;                  // representatives, ascending
  std::vector<double> norm;                 // normalization per representative
  std::vector<State> full;                  // every kept state, ascending
  std::vector<size_t> rep;                  // its representative
  std::vector<std::complex<double>> phase;  // phase of the translation to it
};

inline std::pair<size_t, std::complex<double>> index_and_phase(const BlochBasis& b, const State& x) {
  auto it = std::lower_bound(b.full.begin(), b.full.end(), x);
  if (it == b.full.end() || !(*it == x)) return {b.reps.size(), std::complex<double>(1.)};
  const size_t i = (size_t)(it - b.full.begin());
  return {b.rep[i], b.phase[i]};
}

// n_up < 0: no constraint (every state of the N sites).
inline BlochBasis build_bloch(int N, int n_up, const std::vector<Translation>& trans) {
  std::vector<State> states;
  if (n_up >= 0) {
    states = build_basis(N, n_up);
  } else {
    for (int u = 0; u <= N; ++u) {
      std::vector<State> s = build_basis(N, u);
      states.insert(states.end(), s.begin(), s.end());
    }
    std::sort(states.begin(), states.end());
  }
  BlochBasis b;
  const double l = std::sqrt(double(N));
  for (const State& idx : states) {
    size_t representative_number = b.reps.size();
    std::complex<double> ph = 1., phase_sum = 1.;
    for (size_t t = 1; t < trans.size(); ++t) {
      State tr(idx.size());
      for (size_t i = 0; i < trans[t].perm.size(); ++i) tr[trans[t].perm[i]] = idx[i];
      if (tr > idx) continue;
      if (tr == idx) {
        phase_sum += trans[t].phase;
        continue;
      }
      const size_t i = index_and_phase(b, tr).first;
      if (i < b.reps.size() && b.reps[i] == tr) {
        representative_number = i;
        ph = trans[t].phase;
        break;
      }
    }
    bool storeit = true;
    if (representative_number == b.reps.size()) {
      const double w = std::sqrt(std::abs(phase_sum));
      if (w < 0.01) {
        storeit = false;
      } else {
        b.norm.push_back(w / l);
        b.reps.push_back(idx);
      }
    }
    if (storeit) {
      b.rep.push_back(representative_number);
      b.full.push_back(idx);
      b.phase.push_back(ph);
    }
  }
  return b;
}

// Dense complex H in the Bloch basis (row i, column j), built as the real
// matrix is: explicit zero diagonal, all site terms, then all bonds.
inline std::vector<std::complex<double>> bloch_matrix(const SpinModel& m, const BlochBasis& b) {
  const size_t D = b.reps.size();
  std::vector<std::complex<double>> H(D * D, std::complex<double>(0.));
  auto add = [&](size_t i, const State& ns, double val) {
    const std::pair<size_t, std::complex<double>> jp = index_and_phase(b, ns);
    if (jp.first < D) H[i * D + jp.first] += val * (jp.second * b.norm[jp.first] / b.norm[i]);
  };
  for (int s = 0; s < m.n_sites; ++s)
    for (size_t i = 0; i < D; ++i)
      for (int js = 0; js < 2; ++js) {
        const double val = site_element(m, s, b.reps[i][s], js);
        if (val == 0.0) continue;
        State ns = b.reps[i];
        ns[s] = (short)js;
        add(i, ns, val);
      }
  for (size_t bd = 0; bd < m.bonds.size(); ++bd) {
    const int s1 = m.bonds[bd].s, s2 = m.bonds[bd].t;
    for (size_t i = 0; i < D; ++i)
      for (int j1 = 0; j1 < 2; ++j1)
        for (int j2 = 0; j2 < 2; ++j2) {
          const double val = bond_element(m, (int)bd, b.reps[i][s1], b.reps[i][s2], j1, j2);
          if (val == 0.0) continue;
          State ns = b.reps[i];
          ns[s1] = (short)j1;
          ns[s2] = (short)j2;
          add(i, ns, val);
        }
  }
  return H;
}

// ---------------------------------------------------------------------------
// Start vector: counter-based splitmix64, uniform on [lo, hi)
// ---------------------------------------------------------------------------
// value_i = lo + (hi-lo) * u_i,  u_i = (mix(seed + (i+1)*G) >> 11) * 2^-53,
// G = 0x9e3779b97f4a7c15, mix = the splitmix64 output function (Steele,
// Lea, Flood 2014).  seed 0, i 0 gives the first splitmix64 output of state 0.
inline uint64_t splitmix64_mix(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

inline double start_value(uint64_t seed, uint64_t i, double lo, double hi) {
  uint64_t z = splitmix64_mix(seed + (i + 1) * 0x9e3779b97f4a7c15ULL);
  double u = (double)(z >> 11) * (1.0 / 9007199254740992.0);
  return lo + (hi - lo) * u;
}

// ---------------------------------------------------------------------------
// Symmetric tridiagonal eigensolver (the dstev contract, ietl2lapack.h:67-99)
// ---------------------------------------------------------------------------
// Implicit QL with Wilkinson shifts (Bowdler, Martin, Reinsch, Wilkinson,
// "The QR and QL algorithms for symmetric matrices", Numer. Math. 11, 1968).
// d[0..n) diagonal, e[0..n-1) off-diagonal.  On return d holds the ascending
// eigenvalues and, when z != nullptr, z (column-major n x n, leading dim n)
// the orthonormal eigenvectors.  Returns 0, or >0 when an eigenvalue failed
// to converge (the dstev info convention).
inline int tridiag_eig(int n, double* d, const double* e_in, double* z) {
  if (n <= 0) return 0;
  std::vector<double> e(n, 0.0);
  for (int i = 0; i + 1 < n; ++i) e[i] = e_in[i];
  if (z) {
    for (int i = 0; i < n * n; ++i) z[i] = 0.0;
    for (int i = 0; i < n; ++i) z[i + (size_t)i * n] = 1.0;
  }
  const double eps = std::numeric_limits<double>::epsilon();
  for (int l = 0; l < n; ++l) {
    int iter = 0;
    while (true) {
      int m = l;
      for (; m < n - 1; ++m) {
        double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
        if (std::fabs(e[m]) <= eps * dd) break;
      }
      if (m == l) break;
      if (++iter > 60) return l + 1;
      // Wilkinson shift from the leading 2x2 block at l
      double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
      double r = std::hypot(g, 1.0);
      g = d[m] - d[l] + e[l] / (g + (g >= 0 ? r : -r));
      double s = 1.0, c = 1.0, p = 0.0;
      int i = m - 1;
      bool deflated = false;
      for (; i >= l; --i) {
        double f = s * e[i];
        double b = c * e[i];
        r = std::hypot(f, g);
        e[i + 1] = r;
        if (r == 0.0) {
          d[i + 1] -= p;
          e[m] = 0.0;
          deflated = true;
          break;
        }
        s = f / r;
        c = g / r;
        g = d[i + 1] - p;
        r = (d[i] - g) * s + 2.0 * c * b;
        p = s * r;
        d[i + 1] = g + p;
        g = c * r - b;
        if (z) {
          double* zi = z + (size_t)i * n;
          double* zi1 = z + (size_t)(i + 1) * n;
          for (int k = 0; k < n; ++k) {
            double t = zi1[k];
            zi1[k] = s * zi[k] + c * t;
            zi[k] = c * zi[k] - s * t;
          }
        }
      }
      if (deflated) continue;
      d[l] -= p;
      e[l] = g;
      e[m] = 0.0;
    }
  }
  // ascending order (selection sort keeps it simple and stable enough)
  for (int i = 0; i < n - 1; ++i) {
    int k = i;
    for (int j = i + 1; j < n; ++j)
      if (d[j] < d[k]) k = j;
    if (k != i) {
      std::swap(d[i], d[k]);
      if (z)
        for (int r = 0; r < n; ++r) std::swap(z[r + (size_t)i * n], z[r + (size_t)k * n]);
    }
  }
  return 0;
}

// ---------------------------------------------------------------------------
// Tmatrix (tmatrix.h:24-228)
// ---------------------------------------------------------------------------

struct Tmatrix {
  std::vector<double> alpha, beta;
  double error_tol = std::pow(std::numeric_limits<double>::epsilon(), 2. / 3.);
  double alpha_max = 0, beta_max = 0, beta_min = 0;
  // results of the last compute()
  double thold = 0, multol = 0;
  std::vector<double> err, err_noghost, eig_distinct, eig_noghost;
  std::vector<int> mult, mult_noghost;
  bool computed = false, computed_err = false;

  void push_back(double a, double b) {
    computed = false;
    computed_err = false;
    alpha.push_back(a);
    beta.push_back(b);
    if (alpha.size() == 1) {
      alpha_max = a;
      beta_min = beta_max = b;
    } else {
      if (a > alpha_max) alpha_max = a;
      if (b > beta_max) beta_max = b;
      if (b < beta_min) beta_min = b;
    }
  }

  void compute(bool do_err) {
    err.clear();
    eig_distinct.clear();
    mult.clear();
    err_noghost.clear();
    eig_noghost.clear();
    mult_noghost.clear();
    computed = true;
    computed_err = do_err;
    const int n = (int)alpha.size();
    std::vector<double> eval(alpha.begin(), alpha.end());
    std::vector<double> z;
    if (do_err) z.resize((size_t)n * n);
    int info = tridiag_eig(n, eval.data(), beta.data(), do_err ? z.data() : nullptr);
    if (info > 0) throw std::runtime_error("LAPACK error, stev function failed.");
    const double eps = std::numeric_limits<double>::epsilon();
    multol = std::max(alpha_max, beta_max) * 2 * eps * (10000 + n);
    thold = std::max(eval[0], eval[n - 1]);  // signed maximum, as tmatrix.h:149
    thold = std::max(error_tol * thold, 5 * multol);
    const double bn = beta.back();
    double temp = eval[0];
    eig_distinct.push_back(eval[0]);
    int multiple = 1;
    for (int i = 1; i < n; ++i) {
      if ((eval[i] - temp) > thold) {
        eig_distinct.push_back(eval[i]);
        temp = eval[i];
        mult.push_back(multiple);
        if (do_err) {
          if (multiple > 1)
            err.push_back(0.);
          else
            err.push_back(std::fabs(bn * z[(n - 1) + (size_t)(i - 1) * n]));
        }
        multiple = 1;
      } else {
        multiple++;
      }
    }
    mult.push_back(multiple);
    if (do_err) {
      if (multiple > 1)
        err.push_back(0.);
      else
        err.push_back(std::fabs(bn * z[(n - 1) + (size_t)(n - 1) * n]));
    }
    // ghost test against T with its first row and column removed (:189-217)
    std::vector<double> ag(alpha.begin() + 1, alpha.end());
    std::vector<double> bg(beta.begin() + 1, beta.end());
    info = tridiag_eig(n - 1, ag.data(), bg.data(), nullptr);
    if (info > 0) throw std::runtime_error("LAPACK error, stev function failed.");
    int t2 = 0;
    for (size_t i = 0; i < eig_distinct.size(); ++i) {
      double k = eig_distinct[i];
      if (mult[i] == 1) {
        for (int j = t2; j < n - 2; j++, t2++) {
          if ((ag[j] - k) >= multol) break;
          if (std::fabs(k - ag[j]) < multol) {
            mult[i] = 0;
            if (do_err) err[i] = 0;
            t2++;
            break;
          }
        }
      }
    }
    for (size_t i = 0; i < eig_distinct.size(); ++i) {
      if (mult[i] != 0) {
        eig_noghost.push_back(eig_distinct[i]);
        mult_noghost.push_back(mult[i]);
        if (do_err) err_noghost.push_back(err[i]);
      }
    }
  }

  const std::vector<double>& eigenvalues() {
    if (!computed) compute(false);
    return eig_noghost;
  }
  const std::vector<double>& errors() {
    if (!computed || !computed_err) compute(true);
    return err_noghost;
  }
};

// ---------------------------------------------------------------------------
// Iteration controls (iteration.h:67-134, 209-216)
// ---------------------------------------------------------------------------

struct Iteration {
  bool fixed = false;      // fixed_lanczos_iteration when true
  unsigned max_iter = 0;
  unsigned nev = 1;
  double rtol = 100. * std::numeric_limits<double>::epsilon();
  double atol = 100. * std::numeric_limits<double>::epsilon();
  unsigned i = 0;
  int error = 0;

  bool converged(Tmatrix& T) const {
    if (fixed) return false;
    if (i > 1) {
      const std::vector<double>& errs = T.errors();
      const std::vector<double>& vals = T.eigenvalues();
      if (vals.size() < nev) return false;
      for (unsigned k = 0; k < nev; ++k)
        if (errs[k] > std::max(atol, rtol * std::fabs(vals[k]))) return false;
      return true;
    }
    return false;
  }
  bool finished(Tmatrix& T) {
    if (converged(T)) return true;
    if (i < max_iter) return false;
    error = 1;  // "maximum number of iterations exceeded" (not an exception)
    return true;
  }
};

// ---------------------------------------------------------------------------
// Lanczos (lanczos.h:341-406), FP64 vectors
// ---------------------------------------------------------------------------

// Wall seconds per phase of a solve (for the CPU baseline timing): basis
// build, matrix build, the H.v products, the vector operations of the
// recurrence, and the T-matrix analysis of the stopping test.
struct Phases {
  double basis = 0, matrix = 0, hv = 0, vector = 0, tmatrix = 0;
};
typedef std::chrono::steady_clock phase_clock;
inline double since(phase_clock::time_point t0) {
  return std::chrono::duration<double>(phase_clock::now() - t0).count();
}

struct Lanczos {
  const MapMatrix* H;
  size_t D;
  std::vector<double> startvector, vec2, vec3;
  Tmatrix T;
  unsigned n = 0;
  Phases* ph = nullptr;  // optional phase timers

  void timed_spmv(const double* x, double* y) {
    const auto t0 = phase_clock::now();
    spmv(*H, x, y, D);
    if (ph) ph->hv += since(t0);
  }

  Lanczos(const MapMatrix& h, size_t dim) : H(&h), D(dim), startvector(dim, 0.0), vec2(dim, 0.0), vec3(dim, 0.0) {}

  static double dot(const std::vector<double>& x, const std::vector<double>& y) {
    double t = 0.0;
    for (size_t i = 0; i < x.size(); ++i) t += x[i] * y[i];
    return t;
  }
  static double two_norm(const std::vector<double>& x) {
    double t = 0.0;
    for (size_t i = 0; i < x.size(); ++i) t += x[i] * x[i];
    return std::sqrt(t);
  }

  // make_first_step (:377-390); v0 is the generated start vector
  std::pair<double, double> make_first_step(const std::vector<double>& v0) {
    startvector = v0;
    double nrm = two_norm(startvector);
    for (auto& v : startvector) v /= nrm;
    timed_spmv(startvector.data(), vec2.data());
    double a = dot(startvector, vec2);
    for (size_t i = 0; i < D; ++i) vec2[i] -= a * startvector[i];
    double b = two_norm(vec2);
    for (auto& v : vec2) v /= b;
    return {a, b};
  }

  // make_step (:392-406)
  std::pair<double, double> make_step(unsigned j) {
    double b = T.beta[j - 1];
    timed_spmv(vec2.data(), vec3.data());
    double a = dot(vec2, vec3);
    for (size_t i = 0; i < D; ++i) vec3[i] -= a * vec2[i];
    for (size_t i = 0; i < D; ++i) vec3[i] -= b * startvector[i];
    b = two_norm(vec3);
    for (auto& v : vec3) v /= b;
    std::swap(vec2, startvector);
    std::swap(vec3, vec2);
    return {a, b};
  }

  // generate_tmatrix(iter, gen) then generate_tmatrix(iter) (:341-372)
  void calculate_eigenvalues(Iteration& it, const std::vector<double>& v0) {
    const auto t_all = phase_clock::now();
    double t_test = 0.0, t_hv0 = ph ? ph->hv : 0.0;
    if (T.alpha.empty()) {
      auto ab = make_first_step(v0);
      T.push_back(ab.first, ab.second);
      n = 1;
    }
    for (unsigned j = 0; j < n; ++j) ++it.i;
    for (;;) {
      const auto t0 = phase_clock::now();
      const bool done = it.finished(T);
      t_test += since(t0);
      if (done) break;
      auto ab = make_step(n);
      if (n == T.alpha.size()) T.push_back(ab.first, ab.second);
      ++it.i;
      ++n;
    }
    if (ph) {
      ph->tmatrix += t_test;
      ph->vector += since(t_all) - t_test - (ph->hv - t_hv0);
    }
  }
};

// ---------------------------------------------------------------------------
// sparsediag driver semantics (sparsediag.h:59-149)
// ---------------------------------------------------------------------------

struct SubspaceResult {
  size_t dim = 0;
  std::vector<double> ev;
  std::vector<double> alpha, beta;
  unsigned iterations = 0;
  int error_code = 0;
};

// do_subspace: D==0 -> nothing, D==1 -> H(0,0), else nlowest Lanczos with
// max_iter default min(10*D, 1000) and nev default 1.
inline SubspaceResult do_subspace(const SpinModel& m, int n_up, int max_iter, int nev,
                                  uint64_t seed, bool fixed_iterations = false, Phases* ph = nullptr) {
  SubspaceResult r;
  auto t0 = phase_clock::now();
  std::vector<State> states = build_basis(m.n_sites, n_up);
  if (ph) ph->basis += since(t0);
  r.dim = states.size();
  if (r.dim == 0) return r;
  t0 = phase_clock::now();
  MapMatrix H = build_matrix(m, states);
  if (ph) ph->matrix += since(t0);
  if (r.dim == 1) {
    r.ev.push_back(H[0][0]);
    return r;
  }
  if (max_iter <= 0) max_iter = (int)std::min<size_t>(10 * r.dim, 1000);
  if (nev <= 0) nev = 1;
  Iteration it;
  it.fixed = fixed_iterations;
  it.max_iter = (unsigned)max_iter;
  it.nev = (unsigned)nev;
  Lanczos lz(H, r.dim);
  lz.ph = ph;
  std::vector<double> v0(r.dim);
  for (size_t i = 0; i < r.dim; ++i) v0[i] = start_value(seed, i, -1.0, 1.0);
  lz.calculate_eigenvalues(it, v0);
  const std::vector<double>& ev = lz.T.eigenvalues();
  size_t nn = std::min<size_t>((size_t)nev, ev.size());
  r.ev.assign(ev.begin(), ev.begin() + nn);
  r.alpha = lz.T.alpha;
  r.beta = lz.T.beta;
  r.iterations = it.i;
  r.error_code = it.error;
  return r;
}

// Ground-state energy and gap over sector results (sparsediag.h:63-86).
inline std::pair<double, double> ground_and_gap(const std::vector<std::vector<double>>& sectors) {
  const double big = std::numeric_limits<double>::max();
  double min_val = big, exc_val = big;
  for (const auto& v : sectors) {
    if (v.size() >= 1) {
      if (v[0] < min_val) {
        exc_val = min_val;
        min_val = v[0];
        if (v.size() >= 2 && v[1] < exc_val) exc_val = v[1];
      } else if (v[0] < exc_val) {
        exc_val = v[0];
      }
    }
  }
  return {min_val, exc_val == big ? big : exc_val - min_val};
}

}  // namespace edo
