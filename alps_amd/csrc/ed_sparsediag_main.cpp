// alps_ed_sparsediag -- command-line driver: runs every task of a sparsediag
// input on the GPU and prints one JSON document with the per-sector
// eigenvalues, the ground-state energy and the gap of each task.
//   usage: alps_ed_sparsediag [--describe] [--write-xml] <input>
// The input is a parameter file, a task XML file (<SIMULATION> with
// <PARAMETERS>) or a job XML file (<JOB>, read as the reference's
// MasterScheduler::parse_job_file does, src/alps/scheduler/
// master_scheduler.C:106-168: an optional <OUTPUT>, then <TASK status=...>
// elements with an <INPUT file=...> and an optional <OUTPUT file=...>, file
// names relative to the job file).  Tasks of a job with status "finished"
// are not run.
// --describe resolves lattice, couplings and sectors without solving.
// --write-xml also writes each task's result document: for a job, to the
// task's <OUTPUT> file (its input file when it names none, as the
// reference); otherwise to <file stem>.task<i>.out.xml (i from 1), the name
// the reference's scheduler gives task result files.
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>
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

static bool read_file(const std::string& path, std::string& text) {
  std::ifstream in(path);
  if (!in) return false;
  std::stringstream ss;
  ss << in.rdbuf();
  text = ss.str();
  return true;
}

// The value of attribute `name` in the tag text `tag` ("" when absent);
// white space may surround the '='.
static std::string attribute(const std::string& tag, const std::string& name) {
  for (size_t p = tag.find(name); p != std::string::npos; p = tag.find(name, p + 1)) {
    if (p > 0 && !std::isspace((unsigned char)tag[p - 1])) continue;
    size_t q = p + name.size();
    while (q < tag.size() && std::isspace((unsigned char)tag[q])) ++q;
    if (q >= tag.size() || tag[q] != '=') continue;
    ++q;
    while (q < tag.size() && std::isspace((unsigned char)tag[q])) ++q;
    if (q >= tag.size() || (tag[q] != '"' && tag[q] != '\'')) continue;
    const size_t e = tag.find(tag[q], q + 1);
    if (e == std::string::npos) return "";
    return tag.substr(q + 1, e - q - 1);
  }
  return "";
}

// The tags of an XML text in order ("NAME attr=..." or "/NAME"), skipping
// the declaration, processing instructions and comments.
static std::vector<std::string> xml_tags(const std::string& t) {
  std::vector<std::string> tags;
  for (size_t p = t.find('<'); p != std::string::npos; p = t.find('<', p + 1)) {
    if (t.compare(p, 4, "<!--") == 0) {
      p = t.find("-->", p);
      if (p == std::string::npos) break;
      continue;
    }
    const size_t e = t.find('>', p);
    if (e == std::string::npos) break;
    if (t[p + 1] != '?' && t[p + 1] != '!') {
      std::string tag = t.substr(p + 1, e - p - 1);
      if (!tag.empty() && tag.back() == '/') tag.pop_back();
      tags.push_back(tag);
    }
    p = e;
  }
  return tags;
}

static std::string tag_name(const std::string& tag) {
  size_t e = 0;
  while (e < tag.size() && !std::isspace((unsigned char)tag[e])) ++e;
  return tag.substr(0, e);
}

struct JobTask {
  std::string in, out;  // paths
  bool finished = false;
};

static std::string join_path(const std::string& dir, const std::string& f) {
  if (f.empty() || f[0] == '/' || dir.empty()) return f;
  return dir + "/" + f;
}

// Tasks of a job file; throws std::runtime_error on a malformed job.
static std::vector<JobTask> parse_job(const std::string& text, const std::string& dir) {
  const std::vector<std::string> tags = xml_tags(text);
  size_t i = 0;
  auto name = [&]() { return i < tags.size() ? tag_name(tags[i]) : std::string(); };
  if (name() != "JOB") throw std::runtime_error("missing <JOB> element in jobfile");
  ++i;
  if (name() == "OUTPUT") {
    ++i;
    if (name() == "/OUTPUT") ++i;
  }
  std::vector<JobTask> tasks;
  while (name() == "TASK") {
    JobTask t;
    const std::string status = attribute(tags[i], "status");
    if (status == "finished") t.finished = true;
    else if (!status.empty() && status != "new" && status != "running")
      throw std::runtime_error("illegal status attribute in <TASK> element in jobfile");
    ++i;
    if (name() != "INPUT") throw std::runtime_error("missing <INPUT> element in jobfile");
    t.in = attribute(tags[i], "file");
    if (t.in.empty()) throw std::runtime_error("missing 'file' attribute in <INPUT> element in jobfile");
    ++i;
    if (name() == "/INPUT") ++i;
    if (name() == "OUTPUT") {
      t.out = attribute(tags[i], "file");
      if (t.out.empty()) throw std::runtime_error("missing 'file' attribute in <OUTPUT> element in jobfile");
      ++i;
      if (name() == "/OUTPUT") ++i;
    }
    if (t.out.empty()) t.out = t.in;
    t.in = join_path(dir, t.in);
    t.out = join_path(dir, t.out);
    if (name() != "/TASK") throw std::runtime_error("missing </TASK> tag in jobfile");
    ++i;
    tasks.push_back(t);
  }
  if (name() != "/JOB") throw std::runtime_error("missing </JOB> tag in jobfile");
  return tasks;
}

// Runs the engine on one input text; the JSON output, or false after
// printing the error.
static bool engine(bool describe, const std::string& text, std::string& out) {
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
    return false;
  }
  out = buf.data();
  return true;
}

static bool write_file(const std::string& name, const std::string& text) {
  std::ofstream f(name);
  f << text;
  if (!f) {
    std::fprintf(stderr, "cannot write %s\n", name.c_str());
    return false;
  }
  return true;
}

// The text between the outermost brackets of the engine's {"tasks": [...]}.
static std::string task_list(const std::string& js) {
  const size_t a = js.find('['), b = js.rfind(']');
  if (a == std::string::npos || b == std::string::npos || b <= a) return "";
  return js.substr(a + 1, b - a - 1);
}

static int run_job(bool describe, bool write_xml, const std::string& path, const std::string& text) {
  const size_t slash = path.find_last_of('/');
  const std::string dir = slash == std::string::npos ? std::string() : path.substr(0, slash);
  std::vector<JobTask> tasks;
  try {
    tasks = parse_job(text, dir);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  std::string js = "{\"tasks\": [";
  bool first = true;
  for (const JobTask& t : tasks) {
    if (t.finished) continue;
    std::string in, out;
    if (!read_file(t.in, in)) {
      std::fprintf(stderr, "cannot read %s\n", t.in.c_str());
      return 2;
    }
    if (!engine(describe, in, out)) return 1;
    const std::string items = task_list(out);
    if (!items.empty()) {
      js += (first ? "" : ", ") + items;
      first = false;
    }
    if (write_xml && !describe) {
      const std::vector<std::string> docs = xml_members(out);
      if (docs.size() != 1) {
        std::fprintf(stderr, "error: %s holds %zu tasks, a job task holds one\n", t.in.c_str(), docs.size());
        return 1;
      }
      if (!write_file(t.out, docs[0])) return 1;
    }
  }
  std::printf("%s]}\n", js.c_str());
  return 0;
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
    std::fprintf(stderr, "usage: alps_ed_sparsediag [--describe] [--write-xml] <input>\n");
    return 2;
  }
  std::string text;
  if (!read_file(path, text)) {
    std::fprintf(stderr, "cannot read %s\n", path);
    return 2;
  }
  {
    const std::vector<std::string> tags = xml_tags(text);
    size_t b = 0;
    while (b < text.size() && std::isspace((unsigned char)text[b])) ++b;
    if (b < text.size() && text[b] == '<' && !tags.empty() && tag_name(tags[0]) == "JOB")
      return run_job(describe, write_xml, path, text);
  }
  std::string out;
  if (!engine(describe, text, out)) return 1;
  std::printf("%s\n", out.c_str());
  if (write_xml && !describe) {
    std::string stem = path;
    const size_t slash = stem.find_last_of('/');
    const size_t dot = stem.find_last_of('.');
    if (dot != std::string::npos && (slash == std::string::npos || dot > slash)) stem.erase(dot);
    const std::vector<std::string> docs = xml_members(out);
    for (size_t i = 0; i < docs.size(); ++i)
      if (!write_file(stem + ".task" + std::to_string(i + 1) + ".out.xml", docs[i])) return 1;
  }
  return 0;
}
