// ed_sparsediag.hpp -- the sparsediag task around the device solver.
//
// Mirrors, for the spin-1/2 Sz-conserving path:
//   - lattice graphs: make_graph_from_lattice (src/alps/lattice/
//     latticegraph.h:35-171) over the hypercubic finite lattices
//     (hypercubic.h:110-200,253-280) and unit cells of lib/xml/lattices.xml.in;
//   - the "spin" HAMILTONIAN's parameters and per-type couplings
//     (lib/xml/models.xml.in:342-373, '#' substitution model/substitute.h);
//   - DiagMatrix::dostep / build_subspaces (applications/diag/diag.h:85-175):
//     one solve per Sz_total sector, or the given Sz_total;
//   - SparseDiagMatrix::do_subspace (sparsediag.h:93-149) per sector;
//   - the "Ground State Energy" and "Energy Gap" of write_xml_body
//     (sparsediag.h:59-90).
// The momentum loop inside each Sz sector and the Bloch bases are in
// ed_bloch.hpp.
#pragma once

#include <algorithm>
#include <cmath>
#include <limits>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "ed_params.hpp"

namespace alps_ed {

struct LatticeBond {
  int s, t, type;
};

struct LatticeGraph {
  int n = 0;
  std::vector<int> site_type;
  std::vector<LatticeBond> bonds;
  // geometry of the finite lattice (one vertex per cell): extents and
  // boundaries per dimension, each site's cell offset, the basis vectors
  std::vector<int> extent;
  std::vector<bool> periodic;
  std::vector<std::vector<int>> offset;
  std::vector<std::vector<double>> basis;
};

// A finite lattice: extents, boundary per dimension, one-vertex unit cell
// whose edges go from the cell to the cell at `offset`, in listed order.
struct CellEdge {
  int type;
  std::vector<int> offset;
};

inline LatticeGraph hypercubic_graph(const std::vector<int>& extent, const std::vector<bool>& periodic,
                                     const std::vector<CellEdge>& edges) {
  LatticeGraph g;
  const int dim = (int)extent.size();
  int volume = 1;
  for (int e : extent) {
    if (e < 1) throw std::invalid_argument("lattice extent must be >= 1");
    volume *= e;
  }
  g.n = volume;
  g.site_type.assign(volume, 0);
  g.extent = extent;
  g.periodic = periodic;
  g.basis.assign(dim, std::vector<double>(dim, 0.0));
  for (int d = 0; d < dim; ++d) g.basis[d][d] = 1.0;
  std::vector<int> cell(dim, 0);
  for (int c = 0; c < volume; ++c) {
    // cells in odometer order, last dimension fastest; index row-major
    int ci = 0;
    for (int d = 0; d < dim; ++d) ci = ci * extent[d] + cell[d];
    g.offset.push_back(cell);
    for (const CellEdge& e : edges) {
      int ti = 0;
      bool inside = true;
      for (int d = 0; d < dim && inside; ++d) {
        int v = cell[d] + (d < (int)e.offset.size() ? e.offset[d] : 0);
        if (v < 0 || v >= extent[d]) {
          if (!periodic[d]) inside = false;
          v = ((v % extent[d]) + extent[d]) % extent[d];
        }
        ti = ti * extent[d] + v;
      }
      if (inside && ti != ci) g.bonds.push_back({ci, ti, e.type});
    }
    for (int d = dim - 1; d >= 0; --d) {
      if (++cell[d] < extent[d]) break;
      cell[d] = 0;
    }
  }
  return g;
}

// The LATTICEGRAPHs of lib/xml/lattices.xml.in with one-vertex unit cells.
inline LatticeGraph named_lattice_graph(const std::string& name, const Evaluator& ev) {
  if (name == "dimer")  // an open chain of two sites (lattices.xml.in:295-303)
    return hypercubic_graph({2}, {false}, {{0, {1}}});
  const int L = (int)std::lround(ev.param("L"));
  auto get = [&](const char* p, int def) { return ev.defined(p) ? (int)std::lround(ev.param(p)) : def; };
  std::vector<CellEdge> e;
  if (name == "chain lattice" || name == "open chain lattice") {
    e = {{0, {1}}};
    return hypercubic_graph({L}, {name == "chain lattice"}, e);
  }
  if (name == "nnn chain lattice" || name == "nnn open chain lattice") {
    e = {{0, {1}}, {1, {2}}};
    return hypercubic_graph({L}, {name == "nnn chain lattice"}, e);
  }
  if (name == "square lattice" || name == "open square lattice") {
    const int W = get("W", L);
    const bool p = name == "square lattice";
    return hypercubic_graph({L, W}, {p, p}, {{0, {0, 1}}, {0, {1, 0}}});
  }
  if (name == "anisotropic square lattice") {
    const int W = get("W", L);
    return hypercubic_graph({L, W}, {true, true}, {{0, {1, 0}}, {1, {0, 1}}});
  }
  if (name == "ladder" || name == "open ladder") {
    const int W = get("W", 2);
    return hypercubic_graph({L, W}, {name == "ladder", false}, {{0, {1, 0}}, {1, {0, 1}}});
  }
  if (name == "frustrated square lattice") {
    const int W = get("W", L);
    return hypercubic_graph({L, W}, {true, true}, {{0, {0, 1}}, {0, {1, 0}}, {1, {1, 1}}, {1, {1, -1}}});
  }
  if (name == "triangular lattice" || name == "anisotropic triangular lattice") {
    const int W = get("W", L);
    const bool an = name != "triangular lattice";
    LatticeGraph g =
        hypercubic_graph({L, W}, {true, true}, {{0, {0, 1}}, {an ? 1 : 0, {1, 0}}, {an ? 2 : 0, {-1, 1}}});
    g.basis = {{1.0, 0.0}, {0.5, std::sqrt(3.0) / 2.0}};  // lattices.xml.in:45-49
    return g;
  }
  if (name == "simple cubic lattice" || name == "open simple cubic lattice") {
    const int W = get("W", L);
    const int H = get("H", W);
    const bool p = name == "simple cubic lattice";
    return hypercubic_graph({L, W, H}, {p, p, p}, {{0, {1, 0, 0}}, {0, {0, 1, 0}}, {0, {0, 0, 1}}});
  }
  throw std::invalid_argument("unsupported LATTICE '" + name + "'");
}

// Parameter defaults of the spin HAMILTONIAN (models.xml.in:342-359) and of
// its bond and site terms for the types present ('#' replaced by the type).
inline Parameters with_spin_defaults(Parameters p, const LatticeGraph& g) {
  const char* ham[][2] = {{"J0", "0"},   {"J", "J0"},     {"Jz", "J"},     {"Jxy", "J"},     {"Jz0", "Jz"},
                          {"Jxy0", "Jxy"}, {"J1", "0"},    {"J'", "J1"},    {"Jz'", "J'"},    {"Jxy'", "J'"},
                          {"Jz1", "Jz'"}, {"Jxy1", "Jxy'"}, {"h", "0"},     {"Gamma", "0"},   {"D", "0"},
                          {"K", "0"}};
  for (auto& kv : ham)
    if (!p.count(kv[0])) p[kv[0]] = kv[1];
  std::vector<int> btypes, stypes;
  for (const LatticeBond& b : g.bonds)
    if (std::find(btypes.begin(), btypes.end(), b.type) == btypes.end()) btypes.push_back(b.type);
  for (int t : g.site_type)
    if (std::find(stypes.begin(), stypes.end(), t) == stypes.end()) stypes.push_back(t);
  for (int t : btypes) {
    const std::string s = std::to_string(t);
    if (!p.count("J" + s)) p["J" + s] = "0";
    if (!p.count("Jz" + s)) p["Jz" + s] = "J" + s;
    if (!p.count("Jxy" + s)) p["Jxy" + s] = "J" + s;
    if (!p.count("K" + s)) p["K" + s] = "0";
  }
  for (int t : stypes) {
    const std::string s = std::to_string(t);
    if (!p.count("h" + s)) p["h" + s] = "h";
    if (!p.count("Gamma" + s)) p["Gamma" + s] = "Gamma";
    if (!p.count("D" + s)) p["D" + s] = "D";
  }
  return p;
}

// Spin-1/2 model arrays of one task (bonds in edge order).
struct TaskModel {
  int n_sites = 0;
  std::vector<int> s1, s2;
  std::vector<double> jz, jxy, k, h, d;
  LatticeGraph graph;
};

inline TaskModel task_model(const Parameters& in) {
  if (in.count("MODEL") && in.at("MODEL") != "spin")
    throw std::invalid_argument("only MODEL=\"spin\" is supported (got '" + in.at("MODEL") + "')");
  if (!in.count("LATTICE")) throw std::invalid_argument("LATTICE is not defined");
  Evaluator ev0(in);
  if (ev0.defined("local_S") && std::fabs(ev0.param("local_S") - 0.5) > 1e-12)
    throw std::invalid_argument("only local_S=1/2 is supported");
  const LatticeGraph g = named_lattice_graph(in.at("LATTICE"), ev0);
  const Parameters p = with_spin_defaults(in, g);
  Evaluator ev(p);
  TaskModel m;
  m.n_sites = g.n;
  m.graph = g;
  for (const LatticeBond& b : g.bonds) {
    const std::string t = std::to_string(b.type);
    m.s1.push_back(b.s);
    m.s2.push_back(b.t);
    m.jz.push_back(ev.param("Jz" + t));
    m.jxy.push_back(ev.param("Jxy" + t));
    m.k.push_back(ev.param("K" + t));
  }
  for (int s = 0; s < g.n; ++s) {
    const std::string t = std::to_string(g.site_type[s]);
    if (ev.param("Gamma" + t) != 0.0)
      throw std::invalid_argument("a transverse field Gamma breaks Sz conservation and is not supported");
    m.h.push_back(ev.param("h" + t));
    m.d.push_back(ev.param("D" + t));
  }
  return m;
}

// Sectors of the task (DiagMatrix::build_subspaces and dostep,
// applications/diag/diag.h:86-175, with the spin basis' constraint
// Sz = Sz_total of lib/xml/models.xml.in:228-235):
//   - Sz_total defined: that one sector;
//   - else "Sz" among the CONSERVED_QUANTUMNUMBERS (separated by blanks or
//     commas): every sector from Sz = -N/2 to N/2 in steps of one;
//   - else the whole 2^N space as one subspace, which the engine solves
//     sector by sector and merges (whole_space = true).
struct TaskSectors {
  std::vector<int> n_up;
  bool whole_space = false;
};

inline TaskSectors task_sectors(const Parameters& p, int n_sites) {
  TaskSectors ts;
  Evaluator ev(p);
  if (ev.defined("Sz_total")) {
    const double twice = 2.0 * ev.param("Sz_total") + n_sites;
    const long u = std::lround(twice / 2.0);
    if (std::fabs(twice - 2.0 * u) > 1e-9 || u < 0 || u > n_sites)
      throw std::invalid_argument("Sz_total is not reachable on this lattice");
    ts.n_up.push_back((int)u);
    return ts;
  }
  bool sz = false;
  if (p.count("CONSERVED_QUANTUMNUMBERS")) {
    std::string tok;
    for (char c : p.at("CONSERVED_QUANTUMNUMBERS") + " ") {
      if (c == ' ' || c == ',') {
        sz = sz || tok == "Sz";
        tok.clear();
      } else {
        tok += c;
      }
    }
  }
  for (int u = 0; u <= n_sites; ++u) ts.n_up.push_back(u);
  ts.whole_space = !sz;
  return ts;
}

// Lowest `nev` distinct values of the union of sector spectra: the spectrum
// a Lanczos run on the whole space reports (equal values from different
// sectors are one eigenvalue with multiplicity, merged within rel_tol).
inline std::vector<double> merge_spectra(const std::vector<std::vector<double>>& sectors, int nev,
                                         double rel_tol) {
  std::vector<double> all;
  for (const auto& v : sectors) all.insert(all.end(), v.begin(), v.end());
  std::sort(all.begin(), all.end());
  std::vector<double> out;
  for (double x : all) {
    if (!out.empty() && std::fabs(x - out.back()) <= rel_tol * std::max(1.0, std::fabs(x))) continue;
    out.push_back(x);
    if ((int)out.size() == nev) break;
  }
  return out;
}

// Ground-state energy and gap from the sectors' eigenvalue lists
// (write_xml_body, sparsediag.h:63-86); NaN when not defined.
inline void ground_and_gap(const std::vector<std::vector<double>>& sectors, double& ground, double& gap) {
  const double big = std::numeric_limits<double>::max();
  double min_val = big, exc_val = big;
  for (const auto& v : sectors) {
    if (v.empty()) continue;
    if (v[0] < min_val) {
      exc_val = min_val;
      min_val = v[0];
      if (v.size() >= 2 && v[1] < exc_val) exc_val = v[1];
    } else if (v[0] < exc_val) {
      exc_val = v[0];
    }
  }
  ground = min_val == big ? std::nan("") : min_val;
  gap = exc_val == big ? std::nan("") : exc_val - min_val;
}

}  // namespace alps_ed
