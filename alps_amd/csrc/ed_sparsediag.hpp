// ed_sparsediag.hpp -- the sparsediag task around the device solver.
//
// Mirrors, for the spin-1/2 Sz-conserving path:
//   - lattice graphs: make_graph_from_lattice (src/alps/lattice/
//     latticegraph.h:35-171) over the hypercubic finite lattices
//     (hypercubic.h:110-200,253-280) and unit cells of lib/xml/lattices.xml.in;
//   - explicit graphs (GRAPH, LATTICE_LIBRARY): graph_helper::make_graph
//     (graph_helper.h:651-700) and read_graph_xml (graph.h:120-215);
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
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <fstream>
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

inline bool has_named_lattice(const std::string& name) {
  static const char* names[] = {"dimer", "chain lattice", "open chain lattice", "nnn chain lattice",
                                "nnn open chain lattice", "square lattice", "open square lattice",
                                "anisotropic square lattice", "ladder", "open ladder", "frustrated square lattice",
                                "triangular lattice", "anisotropic triangular lattice", "simple cubic lattice",
                                "open simple cubic lattice"};
  for (const char* n : names)
    if (name == n) return true;
  return false;
}

// ---------------------------------------------------------------------------
// GRAPH: explicit graphs of a lattice library
// ---------------------------------------------------------------------------
// One XML tag: name ("/NAME" for a closing tag), attributes, and whether it
// closes itself ("<EDGE .../>").
struct XmlTag {
  std::string name;
  std::vector<std::pair<std::string, std::string>> attr;
  bool single = false;
  std::string get(const std::string& k) const {
    for (const auto& a : attr)
      if (a.first == k) return a.second;
    return "";
  }
};

// The tags of an XML text in order; declarations, comments and text between
// tags are skipped.
inline std::vector<XmlTag> xml_tags(const std::string& s) {
  std::vector<XmlTag> out;
  size_t i = 0;
  while ((i = s.find('<', i)) != std::string::npos) {
    if (s.compare(i, 4, "<!--") == 0) {
      const size_t e = s.find("-->", i + 4);
      if (e == std::string::npos) throw std::invalid_argument("unterminated comment in lattice library");
      i = e + 3;
      continue;
    }
    if (i + 1 < s.size() && (s[i + 1] == '?' || s[i + 1] == '!')) {
      const size_t e = s.find('>', i);
      if (e == std::string::npos) throw std::invalid_argument("unterminated declaration in lattice library");
      i = e + 1;
      continue;
    }
    XmlTag t;
    size_t j = i + 1;
    auto ws = [&] {
      while (j < s.size() && std::isspace((unsigned char)s[j])) ++j;
    };
    auto word = [&] {
      const size_t b = j;
      if (j == i + 1 && j < s.size() && s[j] == '/') ++j;  // the '/' of a closing tag
      while (j < s.size() && !std::isspace((unsigned char)s[j]) && s[j] != '>' && s[j] != '/' && s[j] != '=') ++j;
      return s.substr(b, j - b);
    };
    t.name = word();
    if (t.name.empty()) throw std::invalid_argument("malformed tag in lattice library");
    for (;;) {
      ws();
      if (j >= s.size()) throw std::invalid_argument("unterminated tag <" + t.name + "> in lattice library");
      if (s[j] == '>') {
        ++j;
        break;
      }
      if (s[j] == '/' && j + 1 < s.size() && s[j + 1] == '>') {
        t.single = true;
        j += 2;
        break;
      }
      const std::string k = word();
      ws();
      if (k.empty() || j >= s.size() || s[j] != '=')
        throw std::invalid_argument("malformed attribute in <" + t.name + "> of lattice library");
      ++j;
      ws();
      if (j >= s.size() || (s[j] != '"' && s[j] != '\''))
        throw std::invalid_argument("unquoted attribute '" + k + "' in <" + t.name + ">");
      const char q = s[j++];
      const size_t e = s.find(q, j);
      if (e == std::string::npos) throw std::invalid_argument("unterminated attribute '" + k + "'");
      t.attr.emplace_back(k, s.substr(j, e - j));
      j = e + 1;
    }
    out.push_back(std::move(t));
    i = j;
  }
  return out;
}

inline int xml_int(const std::string& v, const char* what) {
  char* end = nullptr;
  const long x = std::strtol(v.c_str(), &end, 10);
  if (v.empty() || *end != '\0') throw std::invalid_argument(std::string("bad ") + what + " '" + v + "' in <GRAPH>");
  return (int)x;
}

// The <GRAPH> elements of a lattice library, read as read_graph_xml does
// (src/alps/lattice/graph.h:120-215): vertices="N" fixes the vertex count
// (a VERTEX id beyond it is an error); otherwise vertices are added up to
// the largest id used.  VERTEX: id 1-based (default: the next one), type
// (default 0), an optional COORDINATE inside.  EDGE: source and target
// 1-based, type (default 0), in listed order; an edge's end beyond the
// vertex count adds vertices (boost::add_edge on a vector-backed graph).
inline std::vector<std::pair<std::string, LatticeGraph>> library_graphs(const std::string& xml) {
  std::vector<std::pair<std::string, LatticeGraph>> out;
  const std::vector<XmlTag> tags = xml_tags(xml);
  for (size_t i = 0; i < tags.size(); ++i) {
    if (tags[i].name != "GRAPH") continue;
    const XmlTag& gt = tags[i];
    LatticeGraph g;
    const std::string nv = gt.get("vertices");
    const bool fixed = !nv.empty();
    if (fixed) g.n = xml_int(nv, "vertices");
    if (g.n < 0) throw std::invalid_argument("negative vertex count in <GRAPH>");
    g.site_type.assign(g.n, 0);
    auto grow = [&](int n) {
      if (n > g.n) {
        g.n = n;
        g.site_type.resize(n, 0);
      }
    };
    int next_vertex = 0;
    if (!gt.single) {
      for (++i;; ++i) {
        if (i >= tags.size()) throw std::invalid_argument("closing </GRAPH> tag missing");
        const XmlTag& t = tags[i];
        if (t.name == "/GRAPH") break;
        if (t.name == "VERTEX") {
          const std::string ty = t.get("type"), id = t.get("id");
          const int v = id.empty() ? next_vertex++ : xml_int(id, "VERTEX id") - 1;
          if (v < 0) throw std::invalid_argument("VERTEX id must be >= 1");
          if (v >= g.n) {
            if (fixed) throw std::invalid_argument("too many vertices in <GRAPH>");
            grow(v + 1);
          }
          g.site_type[v] = ty.empty() ? 0 : xml_int(ty, "VERTEX type");
          if (!t.single) {
            ++i;
            if (i < tags.size() && tags[i].name == "COORDINATE") {
              if (!tags[i].single) {
                ++i;
                if (i >= tags.size() || tags[i].name != "/COORDINATE")
                  throw std::invalid_argument("closing </COORDINATE> tag missing");
              }
              ++i;
            }
            if (i >= tags.size() || tags[i].name != "/VERTEX")
              throw std::invalid_argument("closing </VERTEX> tag missing");
          }
        } else if (t.name == "EDGE") {
          const int s = xml_int(t.get("source"), "EDGE source") - 1;
          const int d = xml_int(t.get("target"), "EDGE target") - 1;
          const std::string ty = t.get("type");
          if (s < 0 || d < 0) throw std::invalid_argument("EDGE source and target must be >= 1");
          if (s == d) throw std::invalid_argument("an EDGE from a vertex to itself is not a bond");
          if (!t.single) {
            ++i;
            if (i >= tags.size() || tags[i].name != "/EDGE")
              throw std::invalid_argument("Nonempty <EDGE> tag in <GRAPH>");
          }
          grow(std::max(s, d) + 1);
          g.bonds.push_back({s, d, ty.empty() ? 0 : xml_int(ty, "EDGE type")});
        } else {
          throw std::invalid_argument("illegal element <" + t.name + "> in <GRAPH>");
        }
      }
    }
    out.emplace_back(gt.get("name"), std::move(g));
  }
  return out;
}

// The GRAPHs of lib/xml/lattices.xml.in:617-641.
inline std::vector<std::pair<std::string, LatticeGraph>> builtin_graphs() {
  auto make = [](int n, std::vector<int> types, std::vector<LatticeBond> bonds) {
    LatticeGraph g;
    g.n = n;
    g.site_type = types.empty() ? std::vector<int>(n, 0) : types;
    g.bonds = std::move(bonds);
    return g;
  };
  return {{"5-site dimerized", make(5, {}, {{0, 1, 0}, {1, 2, 1}, {2, 3, 0}, {3, 4, 1}})},
          {"4-site mixed",
           make(4, {0, 1, 0, 1}, {{0, 1, 0}, {1, 2, 0}, {2, 3, 0}, {3, 0, 0}, {0, 2, 1}, {1, 3, 1}})},
          {"triangle", make(3, {}, {{0, 1, 0}, {1, 2, 0}, {2, 0, 0}})}};
}

inline std::string read_text_file(const std::string& path, const char* what) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw std::invalid_argument(std::string("Could not find ") + what + " " + path);
  std::ostringstream o;
  o << f.rdbuf();
  return o.str();
}

// The graph of a task (graph_helper::make_graph, src/alps/lattice/
// graph_helper.h:651-700): GRAPH and LATTICE together are an error; a
// LATTICE name that is one of the lattices above builds that lattice;
// otherwise the name (of GRAPH or LATTICE) is looked up among the GRAPHs of
// the library.  The library is the file LATTICE_LIBRARY names (read from
// the path as given), else the GRAPHs of lattices.xml; the lattices above
// stay available with either library.  A GRAPH has no geometry: no
// translations or momenta, and no distance labels.
inline LatticeGraph task_graph(const Parameters& in, const Evaluator& ev) {
  const bool have_graph = in.count("GRAPH") != 0, have_lattice = in.count("LATTICE") != 0;
  if (have_graph && have_lattice) throw std::invalid_argument("both GRAPH and LATTICE were specified");
  if (!have_graph && !have_lattice) throw std::invalid_argument("LATTICE is not defined");
  const std::string name = in.at(have_graph ? "GRAPH" : "LATTICE");
  if (have_lattice && has_named_lattice(name)) return named_lattice_graph(name, ev);
  const auto graphs = in.count("LATTICE_LIBRARY")
                          ? library_graphs(read_text_file(in.at("LATTICE_LIBRARY"), "lattice library file"))
                          : builtin_graphs();
  for (const auto& ng : graphs)
    if (ng.first == name) {
      if (ng.second.n < 1) throw std::invalid_argument("GRAPH '" + name + "' has no vertices");
      return ng.second;
    }
  throw std::invalid_argument("could not find graph/lattice specified in parameters: name was: '" + name + "'");
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
  Evaluator ev0(in);
  if (ev0.defined("local_S") && std::fabs(ev0.param("local_S") - 0.5) > 1e-12)
    throw std::invalid_argument("only local_S=1/2 is supported");
  const LatticeGraph g = task_graph(in, ev0);
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
