// ed_params.hpp -- the sparsediag parameter input.
//
// Grammar (src/alps/parameter/parameterlist.h:20-36, parameter_p.h:29-47 and
// test/parameter/parameterlist.{input,output}):
//   - assignments `name = value`, separated by ';', ',' or newlines;
//   - values are numbers, expressions or quoted strings ("..." or '...');
//   - `{ ... }` opens a task: global assignments seen so far plus the
//     block's own; an empty block is a task too;
//   - `#clear` forgets the global assignments, `#stop` ends the input;
//   - C and C++ comments; `${NAME}` is replaced by the environment variable.
// Values stay text; numeric evaluation (evaluate) understands + - * / ^,
// parentheses, unary signs, numbers, Pi, sqrt/sin/cos/tan/exp/log/abs and
// references to other parameters, as the ALPS expression parser does for
// the couplings of the spin model.
#pragma once

#include <cctype>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <utility>
#include <stdexcept>
#include <string>
#include <vector>

namespace alps_ed {

// Name/value pairs in the order they were first assigned (the order the
// reference's Parameters keeps and prints); assigning again keeps the place.
class Parameters {
 public:
  typedef std::pair<std::string, std::string> value_type;
  typedef std::vector<value_type>::const_iterator const_iterator;

  size_t count(const std::string& k) const { return find(k) != end() ? 1 : 0; }
  const_iterator find(const std::string& k) const {
    for (auto it = items_.begin(); it != items_.end(); ++it)
      if (it->first == k) return it;
    return items_.end();
  }
  const std::string& at(const std::string& k) const {
    auto it = find(k);
    if (it == end()) throw std::out_of_range("parameter " + k + " is not defined");
    return it->second;
  }
  std::string& operator[](const std::string& k) {
    for (auto& kv : items_)
      if (kv.first == k) return kv.second;
    items_.emplace_back(k, std::string());
    return items_.back().second;
  }
  void clear() { items_.clear(); }
  size_t size() const { return items_.size(); }
  const_iterator begin() const { return items_.begin(); }
  const_iterator end() const { return items_.end(); }

 private:
  std::vector<value_type> items_;
};

namespace detail {

inline std::string substitute_env(const std::string& s) {
  std::string out;
  for (size_t i = 0; i < s.size(); ++i) {
    if (s[i] == '$' && i + 1 < s.size() && s[i + 1] == '{') {
      const size_t e = s.find('}', i + 2);
      if (e == std::string::npos) throw std::invalid_argument("unterminated ${ in parameter value");
      const char* v = std::getenv(s.substr(i + 2, e - i - 2).c_str());
      if (v) out += v;
      i = e;
    } else {
      out += s[i];
    }
  }
  return out;
}

inline std::string trim(const std::string& s) {
  size_t a = 0, b = s.size();
  while (a < b && std::isspace((unsigned char)s[a])) ++a;
  while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
  return s.substr(a, b - a);
}

}  // namespace detail

// Parses a parameter file into its list of tasks.
inline std::vector<Parameters> parse_parameter_file(const std::string& text) {
  std::vector<Parameters> tasks;
  Parameters global;
  Parameters* cur = &global;
  Parameters block;
  bool in_block = false;
  size_t i = 0;
  const size_t n = text.size();
  auto skip_space = [&]() {
    while (i < n) {
      if (std::isspace((unsigned char)text[i])) {
        ++i;
      } else if (text.compare(i, 2, "//") == 0) {
        while (i < n && text[i] != '\n') ++i;
      } else if (text.compare(i, 2, "/*") == 0) {
        const size_t e = text.find("*/", i + 2);
        i = e == std::string::npos ? n : e + 2;
      } else {
        break;
      }
    }
  };
  while (true) {
    skip_space();
    if (i >= n) break;
    const char c = text[i];
    if (c == ';' || c == ',') {
      ++i;
      continue;
    }
    if (c == '{') {
      if (in_block) throw std::invalid_argument("nested '{' in parameter file");
      in_block = true;
      block = global;
      cur = &block;
      ++i;
      continue;
    }
    if (c == '}') {
      if (!in_block) throw std::invalid_argument("unmatched '}' in parameter file");
      tasks.push_back(block);
      in_block = false;
      cur = &global;
      ++i;
      continue;
    }
    if (c == '#') {
      size_t e = i + 1;
      while (e < n && std::isalpha((unsigned char)text[e])) ++e;
      const std::string dir = text.substr(i + 1, e - i - 1);
      i = e;
      if (dir == "clear") {
        global.clear();
        continue;
      }
      if (dir == "stop") break;
      throw std::invalid_argument("unknown directive #" + dir);
    }
    // name
    size_t e = i;
    while (e < n && text[e] != '=' && text[e] != '\n' && text[e] != ';' && text[e] != ',' && text[e] != '{' &&
           text[e] != '}')
      ++e;
    if (e >= n || text[e] != '=') throw std::invalid_argument("expected 'name = value' in parameter file");
    const std::string name = detail::trim(text.substr(i, e - i));
    i = e + 1;
    skip_space();
    std::string value;
    if (i < n && (text[i] == '"' || text[i] == '\'')) {
      const char q = text[i];
      const size_t end = text.find(q, i + 1);
      if (end == std::string::npos) throw std::invalid_argument("unterminated string in parameter file");
      value = text.substr(i + 1, end - i - 1);
      i = end + 1;
    } else {
      size_t end = i;
      int depth = 0;
      while (end < n) {
        const char d = text[end];
        if (d == '(') ++depth;
        if (d == ')') --depth;
        if (depth == 0 && (d == ';' || d == ',' || d == '\n' || d == '}' || d == '{')) break;
        if (text.compare(end, 2, "//") == 0 || text.compare(end, 2, "/*") == 0) break;
        ++end;
      }
      value = detail::trim(text.substr(i, end - i));
      i = end;
    }
    (*cur)[name] = detail::substitute_env(value);
  }
  if (in_block) throw std::invalid_argument("unterminated '{' in parameter file");
  return tasks;
}

// Parameters in the reference's XML forms.  A PARAMETER element sets one
// value (its text, trimmed; the predefined entities decoded) with the
// keep-place assignment of Parameters::operator[]
// (src/alps/parameter/parameters.C:168-177, parameter.C:84-101):
//   - a document whose root is <PARAMETERLIST> (ParameterListXMLHandler,
//     parameterlist.C:69-96): PARAMETER children set defaults, and each
//     <PARAMETERS> child is one task: the defaults so far, then its own
//     values;
//   - otherwise the first <PARAMETERS> element is one task: a task file
//     (<SIMULATION> holding <PARAMETERS>, as parameter2xml writes it,
//     tool/parameter2xml.C:120-130) or a task result file, which repeats
//     its parameters in the same form.
// Comments, processing instructions and other elements are skipped.
namespace detail {
struct XmlTag {
  std::string name;  // "/NAME" for an end tag
  std::string attrs;
  bool empty = false;
  size_t end = 0;  // position after '>'
};

// The next tag at or after `pos` (name empty at the end of the text).
inline XmlTag next_xml_tag(const std::string& t, size_t pos) {
  XmlTag tag;
  for (size_t p = t.find('<', pos); p != std::string::npos; p = t.find('<', p + 1)) {
    if (t.compare(p, 4, "<!--") == 0) {
      const size_t e = t.find("-->", p + 4);
      if (e == std::string::npos) throw std::invalid_argument("XML: unterminated comment");
      p = e + 2;
      continue;
    }
    const size_t e = t.find('>', p);
    if (e == std::string::npos) throw std::invalid_argument("XML: unterminated tag");
    if (t[p + 1] == '?' || t[p + 1] == '!') {
      p = e;
      continue;
    }
    std::string body = t.substr(p + 1, e - p - 1);
    if (!body.empty() && body.back() == '/') {
      tag.empty = true;
      body.pop_back();
    }
    size_t n = 0;
    while (n < body.size() && !std::isspace((unsigned char)body[n])) ++n;
    tag.name = body.substr(0, n);
    tag.attrs = body.substr(n);
    tag.end = e + 1;
    return tag;
  }
  tag.end = t.size();
  return tag;
}

inline std::string xml_decode(std::string v) {
  static const char* ent[][2] = {{"&lt;", "<"}, {"&gt;", ">"}, {"&quot;", "\""}, {"&apos;", "'"}, {"&amp;", "&"}};
  std::string o;
  for (size_t i = 0; i < v.size();) {
    bool hit = false;
    if (v[i] == '&')
      for (auto& e : ent)
        if (v.compare(i, std::strlen(e[0]), e[0]) == 0) {
          o += e[1];
          i += std::strlen(e[0]);
          hit = true;
          break;
        }
    if (!hit) o += v[i++];
  }
  return o;
}

inline std::string xml_attribute(const std::string& attrs, const std::string& name) {
  for (size_t p = attrs.find(name); p != std::string::npos; p = attrs.find(name, p + 1)) {
    if (p > 0 && !std::isspace((unsigned char)attrs[p - 1])) continue;
    size_t q = p + name.size();
    while (q < attrs.size() && std::isspace((unsigned char)attrs[q])) ++q;
    if (q >= attrs.size() || attrs[q] != '=') continue;
    ++q;
    while (q < attrs.size() && std::isspace((unsigned char)attrs[q])) ++q;
    if (q >= attrs.size() || (attrs[q] != '"' && attrs[q] != '\'')) continue;
    const size_t e = attrs.find(attrs[q], q + 1);
    if (e == std::string::npos) break;
    return xml_decode(attrs.substr(q + 1, e - q - 1));
  }
  throw std::invalid_argument("XML: name attribute not found in PARAMETER tag");
}

// One PARAMETER element starting with `tag`: sets it in `p`; returns the
// position after it.
inline size_t read_xml_parameter(const std::string& t, const XmlTag& tag, Parameters& p) {
  const std::string name = xml_attribute(tag.attrs, "name");
  if (tag.empty) {
    p[name] = "";
    return tag.end;
  }
  const size_t close = t.find("</PARAMETER>", tag.end);
  if (close == std::string::npos) throw std::invalid_argument("XML: unterminated <PARAMETER>");
  p[name] = xml_decode(trim(t.substr(tag.end, close - tag.end)));
  return close + 12;
}

// The PARAMETER children of the PARAMETERS element opened by `open`, into
// `p`; returns the position after </PARAMETERS>.
inline size_t read_xml_parameters(const std::string& t, const XmlTag& open, Parameters& p) {
  if (open.empty) return open.end;
  size_t pos = open.end;
  while (true) {
    const XmlTag tag = next_xml_tag(t, pos);
    if (tag.name.empty()) throw std::invalid_argument("XML: unterminated <PARAMETERS>");
    if (tag.name == "/PARAMETERS") return tag.end;
    if (tag.name == "PARAMETER") pos = read_xml_parameter(t, tag, p);
    else throw std::invalid_argument("XML: unknown tag in <PARAMETERS>: " + tag.name);
  }
}
}  // namespace detail

inline std::vector<Parameters> parse_parameter_xml(const std::string& text) {
  using namespace detail;
  XmlTag root = next_xml_tag(text, 0);
  if (root.name == "PARAMETERLIST") {
    std::vector<Parameters> list;
    Parameters defaults;
    size_t pos = root.end;
    while (!root.empty) {
      const XmlTag tag = next_xml_tag(text, pos);
      if (tag.name.empty()) throw std::invalid_argument("XML: unterminated <PARAMETERLIST>");
      if (tag.name == "/PARAMETERLIST") break;
      if (tag.name == "PARAMETER") {
        pos = read_xml_parameter(text, tag, defaults);
      } else if (tag.name == "PARAMETERS") {
        Parameters cur = defaults;
        pos = read_xml_parameters(text, tag, cur);
        list.push_back(cur);
      } else {
        throw std::invalid_argument("XML: unknown tag in <PARAMETERLIST>: " + tag.name);
      }
    }
    return list;
  }
  for (XmlTag tag = root; !tag.name.empty(); tag = next_xml_tag(text, tag.end))
    if (tag.name == "PARAMETERS") {
      Parameters p;
      read_xml_parameters(text, tag, p);
      return {p};
    }
  throw std::invalid_argument("XML input without <PARAMETERS> or <PARAMETERLIST>");
}

// The tasks of an input text: an XML document (its first non-blank
// character is '<') as above, anything else a parameter file.
inline std::vector<Parameters> parse_input(const std::string& text) {
  size_t i = 0;
  while (i < text.size() && std::isspace((unsigned char)text[i])) ++i;
  if (i < text.size() && text[i] == '<') return parse_parameter_xml(text);
  return parse_parameter_file(text);
}

// Numeric value of an expression in the context of a parameter set.
class Evaluator {
 public:
  explicit Evaluator(const Parameters& p) : p_(p) {}

  double value(const std::string& expr, int depth = 0) const {
    if (depth > 64) throw std::invalid_argument("parameter definitions are circular");
    Parser ps{expr, 0, this, depth};
    const double v = ps.sum();
    ps.skip();
    if (ps.i != expr.size()) throw std::invalid_argument("cannot evaluate '" + expr + "'");
    return v;
  }
  bool defined(const std::string& name) const { return p_.count(name) != 0; }
  double param(const std::string& name, int depth = 0) const {
    auto it = p_.find(name);
    if (it == p_.end()) throw std::invalid_argument("parameter " + name + " is not defined");
    return value(it->second, depth + 1);
  }
  double param_or(const std::string& name, double def) const { return defined(name) ? param(name) : def; }

 private:
  struct Parser {
    const std::string& s;
    size_t i;
    const Evaluator* ev;
    int depth;
    void skip() {
      while (i < s.size() && std::isspace((unsigned char)s[i])) ++i;
    }
    double sum() {
      double v = product();
      for (;;) {
        skip();
        if (i < s.size() && (s[i] == '+' || s[i] == '-')) {
          const char op = s[i++];
          const double r = product();
          v = op == '+' ? v + r : v - r;
        } else {
          return v;
        }
      }
    }
    double product() {
      double v = power();
      for (;;) {
        skip();
        if (i < s.size() && (s[i] == '*' || s[i] == '/')) {
          const char op = s[i++];
          const double r = power();
          v = op == '*' ? v * r : v / r;
        } else {
          return v;
        }
      }
    }
    double power() {
      const double b = unary();
      skip();
      if (i < s.size() && s[i] == '^') {
        ++i;
        return std::pow(b, power());
      }
      return b;
    }
    double unary() {
      skip();
      if (i < s.size() && (s[i] == '-' || s[i] == '+')) {
        const char op = s[i++];
        const double v = unary();
        return op == '-' ? -v : v;
      }
      return primary();
    }
    double primary() {
      skip();
      if (i >= s.size()) throw std::invalid_argument("unexpected end of expression '" + s + "'");
      if (s[i] == '(') {
        ++i;
        const double v = sum();
        skip();
        if (i >= s.size() || s[i] != ')') throw std::invalid_argument("missing ')' in '" + s + "'");
        ++i;
        return v;
      }
      if (std::isdigit((unsigned char)s[i]) || s[i] == '.') {
        size_t used = 0;
        const double v = std::stod(s.substr(i), &used);
        i += used;
        return v;
      }
      if (std::isalpha((unsigned char)s[i]) || s[i] == '_') {
        size_t e = i;
        while (e < s.size() && (std::isalnum((unsigned char)s[e]) || s[e] == '_' || s[e] == '\'' || s[e] == '#'))
          ++e;
        const std::string name = s.substr(i, e - i);
        i = e;
        skip();
        if (i < s.size() && s[i] == '(') {
          // the disorder functions draw from the DISORDER_SEED generator
          // (src/alps/expression/evaluator.h:196-228) in the order the
          // reference builds its matrix; couplings here are deterministic
          if (name == "random" || name == "gaussian_random" || name == "normal_random" ||
              name == "integer_random")
            throw std::invalid_argument("random couplings (" + name +
                                        "()) are not supported: give each edge of a GRAPH its own type and value");
          ++i;
          const double a = sum();
          skip();
          if (i >= s.size() || s[i] != ')') throw std::invalid_argument("missing ')' in '" + s + "'");
          ++i;
          if (name == "sqrt") return std::sqrt(a);
          if (name == "sin") return std::sin(a);
          if (name == "cos") return std::cos(a);
          if (name == "tan") return std::tan(a);
          if (name == "exp") return std::exp(a);
          if (name == "log") return std::log(a);
          if (name == "abs") return std::fabs(a);
          throw std::invalid_argument("unknown function " + name);
        }
        if ((name == "Pi" || name == "pi") && !ev->defined(name)) return 3.14159265358979323846;
        return ev->param(name, depth);
      }
      throw std::invalid_argument("cannot evaluate '" + s + "'");
    }
  };
  const Parameters& p_;
};

}  // namespace alps_ed
