// ed_measure.hpp -- eigenvector measurements of the sparsediag task for the
// spin-1/2 path, from the device sums of ed_measure.hip.
//
// Mirrors:
//   - MeasurementOperators (src/alps/scheduler/measurement_operators.C:25-66):
//     MEASURE_AVERAGE[name]=op, MEASURE_LOCAL[name]=op,
//     MEASURE_CORRELATIONS[name]="op1:op2" (or "op" for op:op) and
//     MEASURE_STRUCTURE_FACTOR[name]="op1:op2", kept in name order;
//   - DiagMatrix::perform_measurements (applications/diag/diag.h:191-314):
//     averages of sum_i op(i)/N; local values per site (site operator) or
//     per bond in edge order (bond operator); correlations per distance of
//     the lattice, each from the first site pair at that distance (no
//     translation symmetry); structure factors per momentum,
//     sum_{s1,s2} Re(C[s1][s2] conj(e^{i k.r1}) e^{i k.r2}) / N;
//   - the distances, labels and momenta of the hypercubic finite lattices
//     (src/alps/lattice/hypercubic.h:291-348,352-391, latticegraph.h:204-250,
//     lattice.h:163-181, coordinate_traits.h:97-110).
// Operators: the spin-1/2 site operators Sz, Splus, Sminus
// (lib/xml/models.xml.in:24-36) and Sx (:283-285); the bond operators
// exchange, exchange_xy and biquadratic (:291-301).  Other expressions are
// reported as unsupported, not evaluated.
#pragma once

#include <cmath>
#include <cstdio>
#include <map>
#include <set>
#include <sstream>
#include <string>
#include <utility>
#include <vector>

#include "ed_params.hpp"
#include "ed_sparsediag.hpp"

namespace alps_ed {

struct MeasureSpec {
  std::map<std::string, std::string> average, local;
  std::map<std::string, std::pair<std::string, std::string>> correlation, structure_factor;
  bool any() const { return !average.empty() || !local.empty() || !correlation.empty() || !structure_factor.empty(); }
};

// Correlations and structure factors are labelled by lattice distances and
// momenta, which a GRAPH (no geometry) does not have.
inline void require_geometry(const MeasureSpec& m, const LatticeGraph& g) {
  if ((!m.correlation.empty() || !m.structure_factor.empty()) && g.extent.empty())
    throw std::invalid_argument("MEASURE_CORRELATIONS and MEASURE_STRUCTURE_FACTOR need a LATTICE, not a GRAPH");
}

inline MeasureSpec measure_spec(const Parameters& p) {
  MeasureSpec m;
  auto inner = [](const std::string& key, const std::string& prefix, std::string& name) {
    if (key.size() <= prefix.size() + 1 || key.compare(0, prefix.size(), prefix) != 0 || key.back() != ']')
      return false;
    name = key.substr(prefix.size(), key.size() - prefix.size() - 1);
    return true;
  };
  auto pair_of = [](const std::string& v) {
    const size_t c = v.rfind(':');  // "^(.*):(.*)$": the first group is greedy
    if (c == std::string::npos) return std::make_pair(v, v);
    return std::make_pair(v.substr(0, c), v.substr(c + 1));
  };
  for (const auto& kv : p) {
    std::string name;
    if (inner(kv.first, "MEASURE_AVERAGE[", name)) m.average[name] = kv.second;
    else if (inner(kv.first, "MEASURE_LOCAL[", name)) m.local[name] = kv.second;
    else if (inner(kv.first, "MEASURE_CORRELATIONS[", name)) m.correlation[name] = pair_of(kv.second);
    else if (inner(kv.first, "MEASURE_STRUCTURE_FACTOR[", name)) m.structure_factor[name] = pair_of(kv.second);
  }
  return m;
}

// A spin-1/2 site operator as z Sz + p S+ + m S-.
struct SpinOp {
  double z = 0, p = 0, m = 0;
};

inline bool site_operator(const std::string& name, SpinOp& o) {
  o = SpinOp{};
  if (name == "Sz") o.z = 1;
  else if (name == "Splus") o.p = 1;
  else if (name == "Sminus") o.m = 1;
  else if (name == "Sx") o.p = o.m = 0.5;
  else return false;
  return true;
}

inline bool bond_operator(const std::string& name) {
  return name == "exchange" || name == "exchange_xy" || name == "biquadratic";
}

// Device sums of one vector (ed_measure.hip), divided by |v|^2 on use.
struct SpinSums {
  double n2 = 1;
  std::vector<double> sz, szsz;          // [N], [N*N]
  std::map<std::pair<int, int>, double> flip;  // (i, j) -> <S+_i S-_j>
};

// <A_i> / <A_i B_i> / <A_i B_j> and bond operators, normalised.
struct SpinAlgebra {
  const SpinSums& s;
  int N;
  double sz(int i) const { return s.sz[i] / s.n2; }
  double szsz(int i, int j) const { return s.szsz[(size_t)i * N + j] / s.n2; }
  double flip(int i, int j) const {
    auto it = s.flip.find({i, j});
    if (it == s.flip.end()) throw std::logic_error("flip amplitude not computed");
    return it->second / s.n2;
  }
  double one(const SpinOp& a, int i) const { return a.z * sz(i); }
  // same site: diagonal of the 2x2 product (B acts first), the off-diagonal
  // parts change Sz and vanish in a fixed-Sz vector
  double same(const SpinOp& a, const SpinOp& b, int i) const {
    const double up = a.p * b.m + (0.5 * a.z) * (0.5 * b.z);
    const double dn = a.m * b.p + (-0.5 * a.z) * (-0.5 * b.z);
    return up * (0.5 + sz(i)) + dn * (0.5 - sz(i));
  }
  double two(const SpinOp& a, int i, const SpinOp& b, int j) const {
    double v = 0.0;
    if (a.z != 0 && b.z != 0) v += a.z * b.z * szsz(i, j);
    if (a.p != 0 && b.m != 0) v += a.p * b.m * flip(i, j);
    if (a.m != 0 && b.p != 0) v += a.m * b.p * flip(j, i);
    return v;
  }
  double pair(const SpinOp& a, int i, const SpinOp& b, int j) const { return i == j ? same(a, b, i) : two(a, i, b, j); }
  double bond(const std::string& name, int i, int j) const {
    const double xy = 0.5 * (flip(i, j) + flip(j, i));
    if (name == "exchange_xy") return xy;
    const double ex = szsz(i, j) + xy;
    if (name == "exchange") return ex;
    return 3.0 / 16.0 - 0.5 * ex;  // biquadratic: (S.S)^2 = 3/16 - S.S/2 for spin 1/2
  }
};

// ---------------------------------------------------------------------------
// lattice geometry: coordinates, distances, momenta and their labels
// ---------------------------------------------------------------------------
inline std::vector<double> site_coordinate(const LatticeGraph& g, int s) {
  const size_t dim = g.extent.size();
  std::vector<double> c(dim, 0.0);
  for (size_t d = 0; d < dim; ++d)
    for (size_t x = 0; x < dim; ++x) c[x] += g.offset[s][d] * g.basis[d][x];
  return c;
}

inline std::string coordinate_string(const std::vector<double>& c) {
  std::ostringstream o;
  o << "( ";
  for (size_t i = 0; i < c.size(); ++i) {
    if (i) o << ',';
    o << c[i];
  }
  o << " )";
  return o.str();
}

inline size_t num_distances(const LatticeGraph& g) {
  size_t d = 1;
  for (size_t i = 0; i < g.extent.size(); ++i)
    d *= g.periodic[i] ? (size_t)g.extent[i] : (size_t)g.extent[i] * (size_t)g.extent[i];
  return d;
}

inline size_t distance(const LatticeGraph& g, int s1, int s2) {
  size_t d = 0;
  for (size_t i = 0; i < g.extent.size(); ++i) {
    const size_t e = (size_t)g.extent[i];
    const size_t x = (size_t)g.offset[s1][i], y = (size_t)g.offset[s2][i];
    if (g.periodic[i]) d = d * e + (x <= y ? y - x : e + y - x);
    else d = e * (d * e + x) + y;
  }
  return d;
}

// Reciprocal basis b_i with a_j . b_i = 2 pi delta_ij (the RECIPROCALBASIS
// of lattices.xml.in for the lattices built here).
inline std::vector<std::vector<double>> reciprocal_basis(const LatticeGraph& g) {
  const size_t n = g.extent.size();
  std::vector<std::vector<double>> a = g.basis, inv(n, std::vector<double>(n, 0.0));
  for (size_t i = 0; i < n; ++i) inv[i][i] = 1.0;
  for (size_t c = 0; c < n; ++c) {  // Gauss-Jordan, n <= 3
    size_t piv = c;
    for (size_t r = c + 1; r < n; ++r)
      if (std::fabs(a[r][c]) > std::fabs(a[piv][c])) piv = r;
    std::swap(a[c], a[piv]);
    std::swap(inv[c], inv[piv]);
    const double f = a[c][c];
    for (size_t k = 0; k < n; ++k) {
      a[c][k] /= f;
      inv[c][k] /= f;
    }
    for (size_t r = 0; r < n; ++r)
      if (r != c && a[r][c] != 0.0) {
        const double t = a[r][c];
        for (size_t k = 0; k < n; ++k) {
          a[r][k] -= t * a[c][k];
          inv[r][k] -= t * inv[c][k];
        }
      }
  }
  // rows of A are the a_j, so b_i = 2 pi (column i of A^-1)
  std::vector<std::vector<double>> b(n, std::vector<double>(n));
  for (size_t i = 0; i < n; ++i)
    for (size_t x = 0; x < n; ++x) b[i][x] = 2.0 * M_PI * inv[x][i];
  return b;
}

// Momenta in cell order: k_i = 2 pi x_i / l_i with x_i shifted to
// (-l_i/2, l_i/2] (hypercubic.h:372-386), as Cartesian vectors
// sum_i (k_i / 2 pi) b_i (lattice.h:163-181).
inline std::vector<std::vector<double>> momenta(const LatticeGraph& g) {
  const size_t n = g.extent.size();
  const std::vector<std::vector<double>> b = reciprocal_basis(g);
  std::vector<std::vector<double>> ks;
  for (int s = 0; s < g.n; ++s) {
    std::vector<double> k(n);
    for (size_t i = 0; i < n; ++i) {
      const int l = g.extent[i];
      int x = g.offset[s][i];
      if (x > l / 2) x -= l;
      k[i] = 2. * M_PI * double(x) / double(l);
    }
    std::vector<double> v(n, 0.0);
    if (n > 0) {
      for (size_t j = 0; j < n; ++j) v[j] = b[0][j] * (k[0] / (2. * M_PI));
      for (size_t i = 1; i < n; ++i)
        for (size_t j = 0; j < n; ++j) v[j] = v[j] + b[i][j] * k[i] / (2. * M_PI);
    }
    ks.push_back(v);
  }
  return ks;
}

// ---------------------------------------------------------------------------
// evaluation
// ---------------------------------------------------------------------------
struct MeasurePlan {
  std::vector<std::pair<int, int>> flips;    // ordered pairs the device must sum
  std::vector<std::string> unsupported;      // JSON objects
};

inline std::string json_string(const std::string& v) {
  std::string o = "\"";
  for (char c : v) {
    if (c == '"' || c == '\\') {
      o += '\\';
      o += c;
    } else if (c == '\n') {
      o += "\\n";
    } else if ((unsigned char)c < 0x20) {
      char b[8];
      std::snprintf(b, sizeof b, "\\u%04x", (unsigned)(unsigned char)c);
      o += b;
    } else {
      o += c;
    }
  }
  return o + "\"";
}

inline std::string json_number(double v) {
  if (std::isnan(v) || std::isinf(v)) return "null";
  char b[40];
  std::snprintf(b, sizeof b, "%.17g", v);
  return b;
}

// First site pair at each distance, in the loop order of diag.h:236-241.
inline std::vector<std::pair<int, int>> distance_pairs(const LatticeGraph& g) {
  std::vector<std::pair<int, int>> first(num_distances(g), {-1, -1});
  for (int a = 0; a < g.n; ++a)
    for (int b = 0; b < g.n; ++b) {
      const size_t d = distance(g, a, b);
      if (first[d].first < 0) first[d] = {a, b};
    }
  return first;
}

inline MeasurePlan measure_plan(const MeasureSpec& spec, const LatticeGraph& g) {
  MeasurePlan plan;
  std::set<std::pair<int, int>> need;
  auto flips_of = [&](const SpinOp& a, int i, const SpinOp& b, int j) {
    if (i == j) return;
    if (a.p != 0 && b.m != 0) need.insert({i, j});
    if (a.m != 0 && b.p != 0) need.insert({j, i});
  };
  auto bad = [&](const std::string& kind, const std::string& name, const std::string& why) {
    plan.unsupported.push_back("{\"kind\": " + json_string(kind) + ", \"name\": " + json_string(name) +
                               ", \"reason\": " + json_string(why) + "}");
  };
  SpinOp a, b;
  for (const auto& kv : spec.average)
    if (!site_operator(kv.second, a)) bad("average", kv.first, "'" + kv.second + "' is not a spin-1/2 site operator");
  for (const auto& kv : spec.local) {
    if (site_operator(kv.second, a)) continue;
    if (bond_operator(kv.second)) {
      for (const LatticeBond& e : g.bonds) {
        need.insert({e.s, e.t});
        need.insert({e.t, e.s});
      }
    } else {
      bad("local", kv.first, "'" + kv.second + "' is not a spin-1/2 site or bond operator");
    }
  }
  const std::vector<std::pair<int, int>> dp = distance_pairs(g);
  for (const auto& kv : spec.correlation) {
    if (!site_operator(kv.second.first, a) || !site_operator(kv.second.second, b)) {
      bad("correlations", kv.first, "'" + kv.second.first + ":" + kv.second.second + "' needs spin-1/2 site operators");
      continue;
    }
    for (const auto& pr : dp)
      if (pr.first >= 0) flips_of(a, pr.first, b, pr.second);
  }
  for (const auto& kv : spec.structure_factor) {
    if (!site_operator(kv.second.first, a) || !site_operator(kv.second.second, b)) {
      bad("structure_factor", kv.first,
          "'" + kv.second.first + ":" + kv.second.second + "' needs spin-1/2 site operators");
      continue;
    }
    for (int i = 0; i < g.n; ++i)
      for (int j = 0; j < g.n; ++j) flips_of(a, i, b, j);
  }
  plan.flips.assign(need.begin(), need.end());
  return plan;
}

inline std::string json_rows(const std::vector<std::vector<double>>& rows) {
  std::string o = "[";
  for (size_t i = 0; i < rows.size(); ++i) {
    o += i ? ", [" : "[";
    for (size_t j = 0; j < rows[i].size(); ++j) o += (j ? ", " : "") + json_number(rows[i][j]);
    o += "]";
  }
  return o + "]";
}

inline std::string json_labels(const std::vector<std::string>& v) {
  std::string o = "[";
  for (size_t i = 0; i < v.size(); ++i) o += (i ? ", " : "") + json_string(v[i]);
  return o + "]";
}

// Measured values of one sector's eigenvectors (EigenvectorMeasurements,
// measurement_operators.h:74-98): per name, one value or one labelled list
// per eigenvector.
struct MeasureValues {
  struct Series {
    std::vector<std::string> labels;
    std::vector<std::vector<double>> values;  // [eigenvector][label]
  };
  std::map<std::string, std::vector<double>> average;  // [eigenvector]
  std::map<std::string, Series> local, correlation, structure_factor;
};

inline MeasureValues evaluate_measurements(const MeasureSpec& spec, const LatticeGraph& g,
                                           const std::vector<SpinSums>& vecs) {
  const int N = g.n;
  MeasureValues out;
  std::vector<SpinAlgebra> al;
  for (const SpinSums& s : vecs) al.push_back(SpinAlgebra{s, N});
  SpinOp a, b;
  for (const auto& kv : spec.average) {
    if (!site_operator(kv.second, a)) continue;
    std::vector<double>& v = out.average[kv.first];
    for (const SpinAlgebra& x : al) {
      double t = 0.0;
      for (int i = 0; i < N; ++i) t += x.one(a, i);
      v.push_back(t / N);
    }
  }
  std::vector<std::string> site_labels, bond_labels;
  for (int s = 0; s < N; ++s) site_labels.push_back(coordinate_string(site_coordinate(g, s)));
  for (const LatticeBond& e : g.bonds) bond_labels.push_back(site_labels[e.s] + " -- " + site_labels[e.t]);
  for (const auto& kv : spec.local) {
    const bool site = site_operator(kv.second, a);
    if (!site && !bond_operator(kv.second)) continue;
    MeasureValues::Series& ser = out.local[kv.first];
    ser.labels = site ? site_labels : bond_labels;
    for (const SpinAlgebra& x : al) {
      std::vector<double> r;
      if (site)
        for (int i = 0; i < N; ++i) r.push_back(x.one(a, i));
      else
        for (const LatticeBond& e : g.bonds) r.push_back(x.bond(kv.second, e.s, e.t));
      ser.values.push_back(r);
    }
  }
  const std::vector<std::pair<int, int>> dp = distance_pairs(g);
  std::vector<std::string> dist_labels;
  for (const auto& pr : dp)
    dist_labels.push_back(pr.first < 0 ? std::string() : site_labels[pr.first] + " -- " + site_labels[pr.second]);
  for (const auto& kv : spec.correlation) {
    if (!site_operator(kv.second.first, a) || !site_operator(kv.second.second, b)) continue;
    MeasureValues::Series& ser = out.correlation[kv.first];
    ser.labels = dist_labels;
    for (const SpinAlgebra& x : al) {
      std::vector<double> r;
      for (const auto& pr : dp) r.push_back(pr.first < 0 ? 0.0 : x.pair(a, pr.first, b, pr.second));
      ser.values.push_back(r);
    }
  }
  if (!spec.structure_factor.empty()) {
    const std::vector<std::vector<double>> ks = momenta(g);
    std::vector<std::vector<double>> coord;
    for (int s = 0; s < N; ++s) coord.push_back(site_coordinate(g, s));
    std::vector<std::string> k_labels;
    for (const auto& k : ks) k_labels.push_back(coordinate_string(k));
    for (const auto& kv : spec.structure_factor) {
      if (!site_operator(kv.second.first, a) || !site_operator(kv.second.second, b)) continue;
      MeasureValues::Series& ser = out.structure_factor[kv.first];
      ser.labels = k_labels;
      for (const SpinAlgebra& x : al) {
        std::vector<double> C((size_t)N * N);
        for (int i = 0; i < N; ++i)
          for (int j = 0; j < N; ++j) C[(size_t)i * N + j] = x.pair(a, i, b, j);
        std::vector<double> r;
        for (const auto& k : ks) {
          double t = 0.0;
          for (int i = 0; i < N; ++i)
            for (int j = 0; j < N; ++j) {
              double p1 = 0.0, p2 = 0.0;
              for (size_t d = 0; d < k.size(); ++d) {
                p1 += k[d] * coord[i][d];
                p2 += k[d] * coord[j][d];
              }
              // Re(C conj(e^{i p1}) e^{i p2}) for real C
              t += C[(size_t)i * N + j] * (std::cos(p1) * std::cos(p2) + std::sin(p1) * std::sin(p2)) / N;
            }
          r.push_back(t);
        }
        ser.values.push_back(r);
      }
    }
  }
  return out;
}

// The measurements of one sector as a JSON object.
inline std::string measurements_json(const MeasureValues& mv, const MeasurePlan& plan) {
  auto series = [](const std::map<std::string, MeasureValues::Series>& m) {
    std::string o;
    for (const auto& kv : m)
      o += (o.empty() ? "" : ", ") + json_string(kv.first) + ": {\"labels\": " + json_labels(kv.second.labels) +
           ", \"values\": " + json_rows(kv.second.values) + "}";
    return "{" + o + "}";
  };
  std::string avg;
  for (const auto& kv : mv.average) {
    std::string list = "[";
    for (size_t i = 0; i < kv.second.size(); ++i) list += (i ? ", " : "") + json_number(kv.second[i]);
    avg += (avg.empty() ? "" : ", ") + json_string(kv.first) + ": " + list + "]";
  }
  std::string un = "[";
  for (size_t i = 0; i < plan.unsupported.size(); ++i) un += (i ? ", " : "") + plan.unsupported[i];
  un += "]";
  return "{\"average\": {" + avg + "}, \"local\": " + series(mv.local) + ", \"correlations\": " +
         series(mv.correlation) + ", \"structure_factor\": " + series(mv.structure_factor) +
         ", \"unsupported\": " + un + "}";
}

}  // namespace alps_ed
