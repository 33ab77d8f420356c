// Host-only check of include/ietl/{lanczos,tmatrix,iteration,vectorspace}.h:
// reads a CSR matrix, runs ietl::lanczos with std::vector<double> vectors
// and prints alpha/beta and the distinct non-ghost eigenvalues.
//   usage: ietl_host_lanczos <csr file> <max_iter> <nev> <seed> <fixed 0|1> [vectors]
// With "vectors", also the Ritz vectors of the reported eigenvalues
// (lanczos::eigenvectors) with their Info.  With "resume" (fixed runs): the
// run is also made in two halves, the first archived by lanczos::save into
// a byte archive and continued by a fresh object after lanczos::load, and
// "resume_equal 1" says the halves give the uninterrupted alpha/beta bit
// for bit.
// CSR file: D nnz, then D+1 row pointers, nnz columns, nnz values (text).
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iterator>
#include <string>
#include <type_traits>
#include <vector>

#include <ietl/vectorspace.h>

struct csr_matrix {
  std::vector<uint64_t> rowptr;
  std::vector<uint32_t> col;
  std::vector<double> val;
};

namespace ietl {
// y = prod(m, x): per row, ascending columns (interface/ublas.h:57-63)
inline void mult(const csr_matrix& m, const std::vector<double>& x, std::vector<double>& y) {
  for (size_t i = 0; i + 1 < m.rowptr.size(); ++i) {
    double t = 0.0;
    for (uint64_t p = m.rowptr[i]; p < m.rowptr[i + 1]; ++p) t += m.val[p] * x[m.col[p]];
    y[i] = t;
  }
}
}  // namespace ietl

// the matrix binding must be visible before the algorithm, as with the
// reference's interface headers
#include <ietl/lanczos.h>

// byte archives with the << / >> interface lanczos::save / load use
struct byte_oarchive {
  std::string data;
  template <class T>
  typename std::enable_if<std::is_arithmetic<T>::value, byte_oarchive&>::type operator<<(const T& x) {
    data.append(reinterpret_cast<const char*>(&x), sizeof x);
    return *this;
  }
  template <class T>
  byte_oarchive& operator<<(const std::vector<T>& v) {
    *this << (uint64_t)v.size();
    for (const T& e : v) *this << e;
    return *this;
  }
};
struct byte_iarchive {
  const std::string& data;
  size_t at = 0;
  template <class T>
  typename std::enable_if<std::is_arithmetic<T>::value, byte_iarchive&>::type operator>>(T& x) {
    data.copy(reinterpret_cast<char*>(&x), sizeof x, at);
    at += sizeof x;
    return *this;
  }
  template <class T>
  byte_iarchive& operator>>(std::vector<T>& v) {
    uint64_t n = 0;
    *this >> n;
    v.resize(n);
    for (T& e : v) *this >> e;
    return *this;
  }
};

// splitmix64 counter generator, uniform on [-1, 1)
struct start_gen {
  uint64_t seed, i = 0;
  double operator()() {
    uint64_t z = seed + (++i) * 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    z ^= z >> 31;
    return -1.0 + 2.0 * ((double)(z >> 11) * (1.0 / 9007199254740992.0));
  }
};

int main(int argc, char** argv) {
  if (argc < 6) return 2;
  std::ifstream in(argv[1]);
  size_t D, nnz;
  in >> D >> nnz;
  csr_matrix m;
  m.rowptr.resize(D + 1);
  m.col.resize(nnz);
  m.val.resize(nnz);
  for (auto& v : m.rowptr) in >> v;
  for (auto& v : m.col) in >> v;
  for (auto& v : m.val) in >> v;
  const int max_iter = atoi(argv[2]), nev = atoi(argv[3]);
  const uint64_t seed = strtoull(argv[4], nullptr, 10);
  typedef ietl::vectorspace<std::vector<double> > vs_t;
  vs_t vs(D);
  ietl::lanczos<csr_matrix, vs_t> lz(m, vs);
  int iterations = 0, err = 0;
  if (atoi(argv[5])) {
    ietl::fixed_lanczos_iteration<double> it(max_iter);
    lz.calculate_eigenvalues(it, start_gen{seed});
    iterations = it.iterations();
    err = it.error_code();
  } else {
    ietl::lanczos_iteration_nlowest<double> it(max_iter, nev);
    lz.calculate_eigenvalues(it, start_gen{seed});
    iterations = it.iterations();
    err = it.error_code();
  }
  printf("iterations %d error %d\n", iterations, err);
  for (size_t i = 0; i < lz.diagonal().size(); ++i) printf("ab %.17g %.17g\n", lz.diagonal()[i], lz.subdiagonal()[i]);
  const std::vector<double>& ev = lz.eigenvalues();
  for (size_t i = 0; i < ev.size() && i < (size_t)nev; ++i) printf("ev %.17g\n", ev[i]);
  if (argc > 6 && std::string(argv[6]) == "resume") {
    const int half = max_iter / 2;
    ietl::lanczos<csr_matrix, vs_t> a(m, vs);
    ietl::fixed_lanczos_iteration<double> it1(half);
    a.calculate_eigenvalues(it1, start_gen{seed});
    byte_oarchive oa;
    a.save(oa);
    ietl::lanczos<csr_matrix, vs_t> b(m, vs);
    byte_iarchive ia{oa.data};
    b.load(ia);
    ietl::fixed_lanczos_iteration<double> it2(max_iter);
    b.more_eigenvalues(it2);
    const bool same = b.diagonal() == lz.diagonal() && b.subdiagonal() == lz.subdiagonal() &&
                      ia.at == oa.data.size() && (int)a.diagonal().size() == half;
    printf("resume_equal %d archive_bytes %zu\n", same ? 1 : 0, oa.data.size());
  } else if (argc > 6) {
    std::vector<std::vector<double> > vecs;
    ietl::Info<double> info;
    const size_t k = std::min(ev.size(), (size_t)nev);
    lz.eigenvectors(ev.begin(), ev.begin() + k, std::back_inserter(vecs), info, start_gen{seed});
    for (size_t i = 0; i < vecs.size(); ++i) {
      printf("info %d %d %d %d %.17g %.17g\n", info.m1(i), info.m2(i), info.ma(i), (int)info.error_info(i),
             info.eigenvalue(i), info.residual(i));
      printf("vec");
      for (double v : vecs[i]) printf(" %.17g", v);
      printf("\n");
    }
  }
  return 0;
}
