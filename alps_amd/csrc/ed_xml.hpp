// ed_xml.hpp -- the result file of one sparsediag task (XML).
//
// Mirrors the structure the reference writes for a task:
//   - header and root: Task::write_xml_header / write_xml_trailer
//     (src/alps/scheduler/task.C:220-231), then the task's parameters as
//     <PARAMETERS><PARAMETER name=...>value</PARAMETER>...
//     (src/alps/parameter/parameters.h:279, parameter.h:118-125);
//   - SparseDiagMatrix::write_xml_body (applications/diag/sparsediag/
//     sparsediag.h:59-90): <AVERAGES> with the "Ground State Energy" and
//     "Energy Gap" scalars;
//   - DiagTask::write_xml_body (src/alps/scheduler/diag.hpp:127-160): per
//     sector <EIGENVALUES number=n> with its <QUANTUMNUMBER name value/>
//     list and one eigenvalue per line, then per sector <EIGENSTATES> with
//     one <EIGENSTATE number=j> per level holding
//     EigenvectorMeasurements::write_xml_one_vector
//     (src/alps/scheduler/measurement_operators.h:209-266): scalar averages
//     (including "Energy", sparsediag.h:145-147), then the local values,
//     correlations and structure factors as <VECTOR_AVERAGE> lists of
//     <SCALAR_AVERAGE indexvalue=label><MEAN>.
// Numbers are written with 17 significant digits (boost::lexical_cast of a
// double); quantum numbers as half integers ("1", "-1/2").  Indentation is
// two spaces per level.
#pragma once

#include <cstdio>
#include <string>
#include <utility>
#include <vector>

#include "ed_measure.hpp"
#include "ed_params.hpp"

namespace alps_ed {

inline std::string xml_escape(const std::string& v) {
  std::string o;
  for (char c : v) {
    if (c == '&') o += "&amp;";
    else if (c == '<') o += "&lt;";
    else if (c == '>') o += "&gt;";
    else if (c == '"') o += "&quot;";
    else o += c;
  }
  return o;
}

inline std::string xml_number(double v) {
  char b[40];
  std::snprintf(b, sizeof b, "%.17g", v);
  return b;
}

// Sz of a sector with n_up of n sites up, as the half_integer prints it.
inline std::string half_integer_string(int twice) {
  if (twice % 2 == 0) return std::to_string(twice / 2);
  return std::to_string(twice) + "/2";
}

struct XmlSector {
  std::vector<std::pair<std::string, std::string>> qn;  // (name, value)
  std::vector<double> ev;
  const MeasureValues* meas = nullptr;
};

class XmlWriter {
 public:
  void open(const std::string& tag, const std::vector<std::pair<std::string, std::string>>& attrs = {}) {
    line("<" + tag + attributes(attrs) + ">");
    ++depth_;
  }
  void close(const std::string& tag) {
    --depth_;
    line("</" + tag + ">");
  }
  void empty(const std::string& tag, const std::vector<std::pair<std::string, std::string>>& attrs) {
    line("<" + tag + attributes(attrs) + "/>");
  }
  void inline_text(const std::string& tag, const std::vector<std::pair<std::string, std::string>>& attrs,
                   const std::string& text) {
    line("<" + tag + attributes(attrs) + ">" + xml_escape(text) + "</" + tag + ">");
  }
  void text(const std::string& t) { line(xml_escape(t)); }
  void raw(const std::string& t) { out_ += t + "\n"; }
  // <SCALAR_AVERAGE attr=value><MEAN>v</MEAN></SCALAR_AVERAGE> on one line
  void scalar(const std::string& attr, const std::string& value, double v) {
    line("<SCALAR_AVERAGE " + attr + "=\"" + xml_escape(value) + "\"><MEAN>" + xml_number(v) +
         "</MEAN></SCALAR_AVERAGE>");
  }
  const std::string& str() const { return out_; }

 private:
  static std::string attributes(const std::vector<std::pair<std::string, std::string>>& attrs) {
    std::string o;
    for (const auto& a : attrs) o += " " + a.first + "=\"" + xml_escape(a.second) + "\"";
    return o;
  }
  void line(const std::string& t) { out_ += std::string(2 * depth_, ' ') + t + "\n"; }
  std::string out_;
  int depth_ = 0;
};

// One task's result document.  Each sector writes min(NUMBER_EIGENVALUES,
// its number of values) levels; number_eigenvalues <= 0 means all.  The
// <EIGENSTATES> blocks only when `eigenstates` (measurements requested or
// MEASURE_ENERGY, diag.hpp:141).
inline std::string task_xml(const Parameters& p, const std::vector<XmlSector>& sectors, int number_eigenvalues,
                            bool have_ground, double ground, bool have_gap, double gap, bool eigenstates = true) {
  XmlWriter w;
  w.raw("<?xml version=\"1.0\" encoding=\"UTF-8\"?>");
  w.raw("<?xml-stylesheet type=\"text/xsl\" href=\"ALPS.xsl\"?>");
  w.open("SIMULATION", {{"xmlns:xsi", "http://www.w3.org/2001/XMLSchema-instance"},
                        {"xsi:noNamespaceSchemaLocation", "http://xml.comp-phys.org/2002/10/ALPS.xsd"}});
  w.open("PARAMETERS");
  for (const auto& kv : p) w.inline_text("PARAMETER", {{"name", kv.first}}, kv.second);
  w.close("PARAMETERS");
  w.open("AVERAGES");
  if (have_ground) w.scalar("name", "Ground State Energy", ground);
  if (have_gap) w.scalar("name", "Energy Gap", gap);
  w.close("AVERAGES");
  auto count = [&](const XmlSector& s) {
    const size_t n = number_eigenvalues > 0 ? (size_t)number_eigenvalues : s.ev.size();
    return std::to_string(std::min(n, s.ev.size()));
  };
  for (const XmlSector& s : sectors) {
    w.open("EIGENVALUES", {{"number", count(s)}});
    for (const auto& q : s.qn) w.empty("QUANTUMNUMBER", {{"name", q.first}, {"value", q.second}});
    for (size_t j = 0; j < std::stoul(count(s)); ++j) w.text(xml_number(s.ev[j]));
    w.close("EIGENVALUES");
  }
  for (const XmlSector& s : sectors) {
    if (!eigenstates) break;
    const size_t n = std::stoul(count(s));
    w.open("EIGENSTATES", {{"number", count(s)}});
    for (const auto& q : s.qn) w.empty("QUANTUMNUMBER", {{"name", q.first}, {"value", q.second}});
    for (size_t j = 0; j < n; ++j) {
      w.open("EIGENSTATE", {{"number", std::to_string(j)}});
      // averages in name order, "Energy" among them
      std::map<std::string, double> avg;
      avg["Energy"] = s.ev[j];
      if (s.meas)
        for (const auto& kv : s.meas->average)
          if (j < kv.second.size()) avg[kv.first] = kv.second[j];
      for (const auto& kv : avg) w.scalar("name", kv.first, kv.second);
      if (s.meas) {
        for (const auto* group : {&s.meas->local, &s.meas->correlation, &s.meas->structure_factor})
          for (const auto& kv : *group) {
            if (j >= kv.second.values.size()) continue;
            w.open("VECTOR_AVERAGE", {{"name", kv.first}});
            const std::vector<double>& v = kv.second.values[j];
            for (size_t q = 0; q < v.size() && q < kv.second.labels.size(); ++q)
              w.scalar("indexvalue", kv.second.labels[q], v[q]);
            w.close("VECTOR_AVERAGE");
          }
      }
      w.close("EIGENSTATE");
    }
    w.close("EIGENSTATES");
  }
  w.close("SIMULATION");
  return w.str();
}

}  // namespace alps_ed
