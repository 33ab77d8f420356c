// ed_oracle_capi.cpp -- C entry points of the oracle for the Python tests.
//
// TEST INFRASTRUCTURE ONLY (see ed_oracle.hpp).  Loaded by tests/, by the
// smoke check and by the CPU-baseline leg of bench.py, nowhere else.
#include <chrono>
#include <cstring>
#include <thread>

#include "ed_oracle.hpp"

namespace {

edo::SpinModel make_model(int N, int n_bonds, const int* bs, const int* bt, const double* jz,
                          const double* jxy, const double* k, const double* h, const double* d) {
  edo::SpinModel m;
  m.n_sites = N;
  for (int b = 0; b < n_bonds; ++b) {
    m.bonds.push_back({bs[b], bt[b], 0});
    m.jz.push_back(jz ? jz[b] : 0.0);
    m.jxy.push_back(jxy ? jxy[b] : 0.0);
    m.k.push_back(k ? k[b] : 0.0);
  }
  for (int s = 0; s < N; ++s) {
    m.h.push_back(h ? h[s] : 0.0);
    m.d.push_back(d ? d[s] : 0.0);
  }
  return m;
}

}  // namespace

extern "C" {

// Lattice graph by name; bonds_out receives (source, target, type) triples.
// Returns the number of bonds, or -1 on error, or -(needed) when max_bonds is
// too small.
int oracle_lattice(const char* name, int L, int W, int H, int* n_sites, int* bonds_out,
                   int max_bonds) {
  try {
    edo::Graph g = edo::named_lattice(name, L, W, H);
    *n_sites = g.n;
    int nb = (int)g.bonds.size();
    if (nb > max_bonds) return -nb;
    for (int b = 0; b < nb; ++b) {
      bonds_out[3 * b] = g.bonds[b].s;
      bonds_out[3 * b + 1] = g.bonds[b].t;
      bonds_out[3 * b + 2] = g.bonds[b].type;
    }
    return nb;
  } catch (...) {
    return -1;
  }
}

// Fixed-Sz basis, bit-encoded (bit N-1-s = site s up).  Returns D; writes at
// most max states.
long oracle_basis(int N, int n_up, uint32_t* out, long max) {
  std::vector<edo::State> st = edo::build_basis(N, n_up);
  long D = (long)st.size();
  for (long i = 0; i < D && i < max; ++i) out[i] = edo::state_bits(st[i]);
  return D;
}

// Rows [row0, row1) of H in CSR form (columns ascending).  Returns nnz, or
// -(needed) when max_nnz is too small, or -1 on error.
long oracle_build_csr(int N, int n_up, int n_bonds, const int* bs, const int* bt,
                      const double* jz, const double* jxy, const double* k, const double* h,
                      const double* d, long row0, long row1, uint64_t* rowptr, uint32_t* col,
                      double* val, long max_nnz) {
  try {
    edo::SpinModel m = make_model(N, n_bonds, bs, bt, jz, jxy, k, h, d);
    std::vector<edo::State> st = edo::build_basis(N, n_up);
    if (row1 <= 0 || row1 > (long)st.size()) row1 = (long)st.size();
    edo::MapMatrix H = edo::build_matrix(m, st, (size_t)row0, (size_t)row1);
    edo::CSR c = edo::to_csr(H, (size_t)row0, (size_t)row1);
    long nnz = (long)c.col.size();
    if (nnz > max_nnz) return -nnz;
    std::memcpy(rowptr, c.rowptr.data(), c.rowptr.size() * sizeof(uint64_t));
    std::memcpy(col, c.col.data(), c.col.size() * sizeof(uint32_t));
    std::memcpy(val, c.val.data(), c.val.size() * sizeof(double));
    return nnz;
  } catch (...) {
    return -1;
  }
}

// y = H x with the map-of-maps matrix (full rows).
int oracle_spmv(int N, int n_up, int n_bonds, const int* bs, const int* bt, const double* jz,
                const double* jxy, const double* k, const double* h, const double* d,
                const double* x, double* y) {
  try {
    edo::SpinModel m = make_model(N, n_bonds, bs, bt, jz, jxy, k, h, d);
    std::vector<edo::State> st = edo::build_basis(N, n_up);
    edo::MapMatrix H = edo::build_matrix(m, st);
    edo::spmv(H, x, y, st.size());
    return 0;
  } catch (...) {
    return -1;
  }
}

// dstev contract: eigenvalues ascending into eval, eigenvectors into z
// (column-major n x n) when z is not null.
int oracle_tridiag(int n, const double* d, const double* e, double* eval, double* z) {
  for (int i = 0; i < n; ++i) eval[i] = d[i];
  return edo::tridiag_eig(n, eval, e, z);
}

// Tmatrix::compute(true) on a given alpha/beta (beta[n-1] is the next beta).
// Outputs the distinct non-ghost eigenvalues, their errors and multiplicities.
int oracle_tmatrix(int n, const double* alpha, const double* beta, double* eig, double* err,
                   int* mult, int* n_out, double* thold, double* multol) {
  try {
    edo::Tmatrix T;
    for (int i = 0; i < n; ++i) T.push_back(alpha[i], beta[i]);
    T.compute(true);
    int m = (int)T.eig_noghost.size();
    for (int i = 0; i < m; ++i) {
      eig[i] = T.eig_noghost[i];
      err[i] = T.err_noghost[i];
      mult[i] = T.mult_noghost[i];
    }
    *n_out = m;
    *thold = T.thold;
    *multol = T.multol;
    return 0;
  } catch (...) {
    return -1;
  }
}

// write_xml_body's ground-state energy and gap over sector results; counts[i]
// values of sector i follow each other in `values`.  The gap is DBL_MAX when
// undefined, as the reference's sentinel.
int oracle_ground_and_gap(int n_sectors, const int* counts, const double* values, double* ground,
                          double* gap) {
  std::vector<std::vector<double>> sec(n_sectors);
  size_t o = 0;
  for (int i = 0; i < n_sectors; ++i) {
    sec[i].assign(values + o, values + o + counts[i]);
    o += counts[i];
  }
  std::pair<double, double> r = edo::ground_and_gap(sec);
  *ground = r.first;
  *gap = r.second;
  return 0;
}

double oracle_start_value(uint64_t seed, uint64_t i, double lo, double hi) {
  return edo::start_value(seed, i, lo, hi);
}

// Full do_subspace: eigenvalues (ev, up to nev), the alpha/beta sequence (up
// to max_ab entries), iteration count and the iteration error code.
int oracle_sparsediag(int N, int n_up, int n_bonds, const int* bs, const int* bt,
                      const double* jz, const double* jxy, const double* k, const double* h,
                      const double* d, int max_iter, int nev, uint64_t seed, int fixed,
                      double* ev_out, int* n_ev, double* alpha_out, double* beta_out,
                      int max_ab, int* n_ab, int* iterations, int* error_code, long* dim) {
  try {
    edo::SpinModel m = make_model(N, n_bonds, bs, bt, jz, jxy, k, h, d);
    edo::SubspaceResult r = edo::do_subspace(m, n_up, max_iter, nev, seed, fixed != 0);
    *dim = (long)r.dim;
    *n_ev = (int)r.ev.size();
    for (size_t i = 0; i < r.ev.size(); ++i) ev_out[i] = r.ev[i];
    int na = (int)r.alpha.size();
    *n_ab = na;
    for (int i = 0; i < na && i < max_ab; ++i) {
      alpha_out[i] = r.alpha[i];
      beta_out[i] = r.beta[i];
    }
    *iterations = (int)r.iterations;
    *error_code = r.error_code;
    return 0;
  } catch (...) {
    return -1;
  }
}

// do_subspace with wall seconds per phase: phases[0..5) = basis build,
// matrix build, H.v products, vector operations, T-matrix analysis.
int oracle_sparsediag_timed(int N, int n_up, int n_bonds, const int* bs, const int* bt, const double* jz,
                            const double* jxy, const double* k, const double* h, const double* d, int max_iter,
                            int nev, uint64_t seed, int fixed, double* ev_out, int* n_ev, int* iterations,
                            long* dim, double* phases) {
  try {
    edo::SpinModel m = make_model(N, n_bonds, bs, bt, jz, jxy, k, h, d);
    edo::Phases ph;
    edo::SubspaceResult r = edo::do_subspace(m, n_up, max_iter, nev, seed, fixed != 0, &ph);
    *dim = (long)r.dim;
    *n_ev = (int)r.ev.size();
    for (size_t i = 0; i < r.ev.size(); ++i) ev_out[i] = r.ev[i];
    *iterations = (int)r.iterations;
    phases[0] = ph.basis;
    phases[1] = ph.matrix;
    phases[2] = ph.hv;
    phases[3] = ph.vector;
    phases[4] = ph.tmatrix;
    return 0;
  } catch (...) {
    return -1;
  }
}

// CPU baseline of the H.v hot loop: builds rows [0, rows) of H with the
// reference's binary-search lookup into the map-of-maps, then times `reps`
// row-wise products over them.  Returns seconds per product of the sample;
// build_seconds receives the build time.
double oracle_bench_hv(int N, int n_up, int n_bonds, const int* bs, const int* bt,
                       const double* jz, const double* jxy, const double* k, const double* h,
                       const double* d, long rows, int reps, int threads, double* build_seconds,
                       long* sample_nnz) {
  typedef std::chrono::steady_clock clk;
  edo::SpinModel m = make_model(N, n_bonds, bs, bt, jz, jxy, k, h, d);
  std::vector<edo::State> st = edo::build_basis(N, n_up);
  if (rows <= 0 || rows > (long)st.size()) rows = (long)st.size();
  auto t0 = clk::now();
  edo::MapMatrix H = edo::build_matrix(m, st, 0, (size_t)rows);
  auto t1 = clk::now();
  *build_seconds = std::chrono::duration<double>(t1 - t0).count();
  long nnz = 0;
  for (const auto& r : H) nnz += (long)r.second.size();
  *sample_nnz = nnz;
  std::vector<double> x(st.size()), y(st.size());
  for (size_t i = 0; i < st.size(); ++i) x[i] = edo::start_value(1, i, -1.0, 1.0);
  // rows split into `threads` contiguous ranges of the map (one thread
  // each); the median of `reps` products
  const int nt = std::max(1, threads);
  std::vector<edo::MapMatrix::const_iterator> cut;
  {
    size_t i = 0;
    const size_t per = (H.size() + nt - 1) / nt;
    for (auto it = H.begin(); it != H.end(); ++it, ++i)
      if (i % per == 0) cut.push_back(it);
    cut.push_back(H.end());
  }
  std::vector<double> times;
  for (int r = 0; r < reps; ++r) {
    auto a = clk::now();
    auto part = [&](size_t c) {
      for (auto row = cut[c]; row != cut[c + 1]; ++row) {
        double t = 0.0;
        for (const auto& e : row->second) t += e.second * x[e.first];
        y[row->first] = t;
      }
    };
    std::vector<std::thread> pool;
    for (size_t c = 1; c + 1 < cut.size(); ++c) pool.emplace_back(part, c);
    part(0);
    for (auto& t : pool) t.join();
    auto b = clk::now();
    times.push_back(std::chrono::duration<double>(b - a).count());
  }
  std::sort(times.begin(), times.end());
  // keep the result observable
  volatile double sink = y[0];
  (void)sink;
  return times[times.size() / 2];
}

}  // extern "C"

// <v| O |v> of a site term on the fixed-Sz basis of N sites with n_up up
// spins (v in basis order): O = A(i) when j < 0, A(i)*B(i) when j == i,
// A(i)*B(j) otherwise; operators 0 Sz, 1 Splus, 2 Sminus, 3 Sx.
extern "C" int oracle_expect_site_term(int N, int n_up, const double* v, int opA, int i, int opB, int j,
                                       double* out) {
  try {
    std::vector<edo::State> st = edo::build_basis(N, n_up);
    *out = edo::expect_site_term(st, v, opA, i, opB, j);
    return 0;
  } catch (...) {
    return -1;
  }
}

namespace {

std::vector<edo::Translation> read_translations(int N, int n_t, const long* perm, const double* re,
                                                const double* im) {
  std::vector<edo::Translation> t(n_t);
  for (int i = 0; i < n_t; ++i) {
    t[i].phase = std::complex<double>(re[i], im[i]);
    t[i].perm.assign(perm + (size_t)i * N, perm + (size_t)(i + 1) * N);
  }
  return t;
}

}  // namespace

// Momenta of a named lattice (one row of n_dirs values per momentum).
// Returns the number of momenta (-1 on error); *n_dirs = periodic directions.
extern "C" int oracle_translation_momenta(const char* name, int L, int W, int H, int max, double* out,
                                          int* n_dirs) {
  try {
    const edo::Graph g = edo::named_lattice(name, L, W, H);
    const std::vector<std::vector<double>> ks = edo::translation_momenta(g);
    *n_dirs = ks.empty() ? 0 : (int)ks[0].size();
    for (int i = 0; i < (int)ks.size() && i < max; ++i)
      for (int d = 0; d < *n_dirs; ++d) out[i * *n_dirs + d] = ks[i][d];
    return (int)ks.size();
  } catch (...) {
    return -1;
  }
}

// Translations of a named lattice at momentum k: phases and site maps
// (perm[t*N + i] = image of site i).  Returns their number (-1 on error).
extern "C" int oracle_translations(const char* name, int L, int W, int H, int n_k, const double* k, int max,
                                   long* perm, double* re, double* im) {
  try {
    const edo::Graph g = edo::named_lattice(name, L, W, H);
    const std::vector<edo::Translation> t = edo::translations(g, std::vector<double>(k, k + n_k));
    for (int i = 0; i < (int)t.size() && i < max; ++i) {
      re[i] = t[i].phase.real();
      im[i] = t[i].phase.imag();
      for (int s = 0; s < g.n; ++s) perm[(size_t)i * g.n + s] = (long)t[i].perm[s];
    }
    return (int)t.size();
  } catch (...) {
    return -1;
  }
}

// Representatives (bit-encoded) and normalizations of the Bloch basis of N
// sites (n_up < 0: no constraint).  Returns their number (-1 on error).
extern "C" long oracle_bloch_basis(int N, int n_up, int n_t, const long* perm, const double* re, const double* im,
                                   long max, uint32_t* reps, double* norm) {
  try {
    const edo::BlochBasis b = edo::build_bloch(N, n_up, read_translations(N, n_t, perm, re, im));
    for (size_t i = 0; i < b.reps.size() && (long)i < max; ++i) {
      reps[i] = edo::state_bits(b.reps[i]);
      norm[i] = b.norm[i];
    }
    return (long)b.reps.size();
  } catch (...) {
    return -1;
  }
}

// Dense complex H of the model in the Bloch basis (row-major, real and
// imaginary parts).  Returns its dimension (-1 on error; -(dim) if it
// exceeds max_dim).
extern "C" long oracle_bloch_dense(int N, int n_up, int n_bonds, const int* bs, const int* bt, const double* jz,
                                   const double* jxy, const double* k, const double* h, const double* d, int n_t,
                                   const long* perm, const double* re, const double* im, long max_dim,
                                   double* out_re, double* out_im) {
  try {
    const edo::SpinModel m = make_model(N, n_bonds, bs, bt, jz, jxy, k, h, d);
    const edo::BlochBasis b = edo::build_bloch(N, n_up, read_translations(N, n_t, perm, re, im));
    const long D = (long)b.reps.size();
    if (D > max_dim) return -D;
    const std::vector<std::complex<double>> H = edo::bloch_matrix(m, b);
    for (size_t i = 0; i < H.size(); ++i) {
      out_re[i] = H[i].real();
      out_im[i] = H[i].imag();
    }
    return D;
  } catch (...) {
    return -1;
  }
}
