// alps_ed_sparsediag -- command-line driver: runs every task of a sparsediag
// parameter file on the GPU and prints one JSON document with the
// per-sector eigenvalues, the ground-state energy and the gap of each task.
//   usage: alps_ed_sparsediag [--describe] [--write-xml] <parameter file>
// --describe resolves lattice, couplings and sectors without solving.
// --write-xml also writes each task's result document as
// <file stem>.task<i>.out.xml (i from 1), the name the reference's
// scheduler gives task result files.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include <alps_ed.h>

// The values of every "xml" member of the engine's JSON output, unescaped
// (its strings escape only '"', '\\', newlines and \u00XX control codes).
static std::vector<std::string> xml_members(const std::string& js) {
  std::vector<std::string> out;
  const std::string key = "\"xml\": \"";
  for (size_t p = js.find(key); p != std::string::npos; p = js.find(key, p)) {
    p += key.size();
    std::string v;
    while (p < js.size() && js[p] != '"') {
      if (js[p] == '\\' && p + 1 < js.size()) {
        const char c = js[++p];
        if (c == 'n') v += '\n';
        else if (c == 'u' && p + 4 < js.size()) {
          v += (char)std::strtol(js.substr(p + 1, 4).c_str(), nullptr, 16);
          p += 4;
        } else v += c;
      } else {
        v += js[p];
      }
      ++p;
    }
    out.push_back(v);
  }
  return out;
}

int main(int argc, char** argv) {
  bool describe = false, write_xml = false;
  const char* path = nullptr;
  for (int i = 1; i < argc; ++i) {
    if (std::strcmp(argv[i], "--describe") == 0)
      describe = true;
    else if (std::strcmp(argv[i], "--write-xml") == 0)
      write_xml = true;
    else
      path = argv[i];
  }
  if (!path) {
    std::fprintf(stderr, "usage: alps_ed_sparsediag [--describe] [--write-xml] <parameter file>\n");
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
  if (write_xml && !describe) {
    std::string stem = path;
    const size_t slash = stem.find_last_of('/');
    const size_t dot = stem.find_last_of('.');
    if (dot != std::string::npos && (slash == std::string::npos || dot > slash)) stem.erase(dot);
    const std::vector<std::string> docs = xml_members(buf.data());
    for (size_t i = 0; i < docs.size(); ++i) {
      const std::string name = stem + ".task" + std::to_string(i + 1) + ".out.xml";
      std::ofstream f(name);
      f << docs[i];
      if (!f) {
        std::fprintf(stderr, "cannot write %s\n", name.c_str());
        return 1;
      }
    }
  }
  return 0;
}
