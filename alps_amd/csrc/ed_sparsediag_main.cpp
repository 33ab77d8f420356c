// alps_ed_sparsediag -- command-line driver: runs every task of a sparsediag
// parameter file on the GPU and prints one JSON document with the
// per-sector eigenvalues, the ground-state energy and the gap of each task.
//   usage: alps_ed_sparsediag [--describe] <parameter file>
// --describe resolves lattice, couplings and sectors without solving.
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include <alps_ed.h>

int main(int argc, char** argv) {
  bool describe = false;
  const char* path = nullptr;
  for (int i = 1; i < argc; ++i) {
    if (std::strcmp(argv[i], "--describe") == 0)
      describe = true;
    else
      path = argv[i];
  }
  if (!path) {
    std::fprintf(stderr, "usage: alps_ed_sparsediag [--describe] <parameter file>\n");
    return 2;
  }
  std::ifstream in(path);
  if (!in) {
    std::fprintf(stderr, "cannot read %s\n", path);
    return 2;
  }
  std::stringstream ss;
  ss << in.rdbuf();
  const std::string text = ss.str();
  auto run = describe ? alps_ed_describe_parameter_file : alps_ed_run_parameter_file;
  std::vector<char> buf(1 << 20);
  size_t need = 0;
  int st = run(text.c_str(), buf.data(), buf.size(), &need);
  if (st == ALPS_ED_OK && need > buf.size()) {
    // only a very long output needs a second run with the full length
    buf.resize(need);
    st = run(text.c_str(), buf.data(), buf.size(), &need);
  }
  if (st != ALPS_ED_OK) {
    std::fprintf(stderr, "error: %s\n", alps_ed_last_error());
    return 1;
  }
  std::printf("%s\n", buf.data());
  return 0;
}
