// ietl::lanczos<alps_ed::hamiltonian, alps_ed::gpu_vectorspace>: the
// sparsediag Lanczos call (applications/diag/sparsediag/sparsediag.h:104-118)
// with the device operator and device vectors.
//   usage: ietl_gpu_lanczos <L> <n_up> <f32|f64> <max_iter> <nev> <seed> [vectors]
// With "vectors", also the Ritz vectors (lanczos::eigenvectors, replaying the
// device recurrence) with their Info and the first components.
// Heisenberg ring J=1 on L sites; prints iterations and eigenvalues.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iterator>
#include <vector>

#include <alps_ed/ietl_gpu.hpp>
#include <ietl/lanczos.h>

int main(int argc, char** argv) {
  if (argc < 7) return 2;
  const int L = atoi(argv[1]), n_up = atoi(argv[2]);
  const alps_ed_dtype t = strcmp(argv[3], "f64") == 0 ? ALPS_ED_F64 : ALPS_ED_F32;
  const int max_iter = atoi(argv[4]), nev = atoi(argv[5]);
  const uint64_t seed = strtoull(argv[6], nullptr, 10);
  std::vector<int> s1(L), s2(L);
  std::vector<double> j(L, 1.0);
  for (int i = 0; i < L; ++i) {
    s1[i] = i;
    s2[i] = (i + 1) % L;
  }
  alps_ed_model m{L, n_up, L, s1.data(), s2.data(), j.data(), j.data(), nullptr, nullptr, nullptr};
  try {
    alps_ed::hamiltonian H(m);
    alps_ed::gpu_vectorspace vs(H.dimension(), t);
    ietl::lanczos<alps_ed::hamiltonian, alps_ed::gpu_vectorspace> lz(H, vs);
    ietl::lanczos_iteration_nlowest<double> it(max_iter, nev);
    lz.calculate_eigenvalues(it, alps_ed::uniform_start{seed});
    printf("dim %llu iterations %u error %d\n", (unsigned long long)H.dimension(), it.iterations(), it.error_code());
    const std::vector<double>& ev = lz.eigenvalues();
    for (size_t i = 0; i < ev.size() && i < (size_t)nev; ++i) printf("ev %.17g\n", ev[i]);
    std::fflush(stdout);
    if (argc > 7) {
      std::vector<alps_ed::gpu_vector> vecs;
      ietl::Info<double> info;
      const size_t k = std::min(ev.size(), (size_t)nev);
      lz.eigenvectors(ev.begin(), ev.begin() + k, std::back_inserter(vecs), info, alps_ed::uniform_start{seed});
      for (size_t i = 0; i < vecs.size(); ++i) {
        printf("info %d %d %d %d %.17g %.17g %.17g\n", info.m1(i), info.m2(i), info.ma(i), (int)info.error_info(i),
               info.eigenvalue(i), info.residual(i), alps_ed::two_norm(vecs[i]));
        const std::vector<double> h = vecs[i].to_host();
        printf("vec");
        for (size_t q = 0; q < h.size() && q < 4096; ++q) printf(" %.17g", h[q]);
        printf("\n");
      }
      std::fflush(stdout);
    }
  } catch (const std::exception& e) {
    printf("exception %s\n", e.what());
    return 1;
  }
  // every device object of this program is destroyed above; the output is
  // complete, and the process ends without running exit-time teardown
  // (library and runtime destructors), which this check does not cover
  std::fflush(stdout);
  std::_Exit(0);
}
