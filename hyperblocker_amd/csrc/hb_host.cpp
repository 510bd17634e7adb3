// Host front end of the blocking path: rule files -> execution plan, CSV
// tables, hash partitioning and the match writer.  Pure C++ (no device).
//
// Reference counterparts (paths relative to the HyperBlocker tree):
//   rule decoding      core/common/yaml_config.h:19-57 (YAML::convert<Rule>)
//   plan generation    core/components/execution_plan_generator.h:64-155
//   CSV load           core/components/data_mngr.h:46-79 (rapidcsv Document)
//   partitioning       core/components/data_mngr.h:81-222 (DataPartitioning)
//   output             core/components/host_reducer.h:80-90 (WriteMatch)
//
// yaml-cpp, rapidcsv and gflags are not available offline, so the YAML
// subset and the CSV dialect the reference relies on are implemented here.
#include <algorithm>
#include <cerrno>
#include <climits>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <dirent.h>
#include <fstream>
#include <functional>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <string_view>
#include <sys/stat.h>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/hyperblocker_amd.h"

namespace hbrt {
hb_status fail(hb_status st, const char *fmt, ...);
}
using hbrt::fail;

namespace {

// ---- YAML subset -------------------------------------------------------------
// Block mappings and sequences by indentation, flow sequences/mappings,
// plain / single- / double-quoted scalars, multi-line plain scalars,
// comments.  Enough for the rule files of the reference (data/*/rules).

struct YNode {
  enum Kind { kNull, kScalar, kSeq, kMap } kind = kNull;
  std::string scalar;
  std::vector<std::unique_ptr<YNode>> seq;
  std::vector<std::pair<std::string, std::unique_ptr<YNode>>> map;
  const YNode *get(const std::string &k) const {
    if (kind != kMap) return nullptr;
    for (auto &kv : map)
      if (kv.first == k) return kv.second.get();
    return nullptr;
  }
  const YNode *at(size_t i) const {
    if (kind != kSeq || i >= seq.size()) return nullptr;
    return seq[i].get();
  }
  size_t size() const {
    return kind == kSeq ? seq.size() : kind == kMap ? map.size() : 0;
  }
};

struct YLine {
  int indent;
  std::string text;  // without indentation, comment and trailing spaces
  int lineno;
};

class YamlParser {
 public:
  explicit YamlParser(const std::string &src) { split(src); }

  bool parse(std::unique_ptr<YNode> &root, std::string &err) {
    pos_ = 0;
    err_.clear();
    root.reset(new YNode());
    if (lines_.empty()) return true;
    if (!block(lines_[0].indent, *root)) {
      err = err_;
      return false;
    }
    if (pos_ < lines_.size()) {
      err = "line " + std::to_string(lines_[pos_].lineno) + ": unexpected indentation";
      return false;
    }
    return true;
  }

 private:
  static std::string rstrip(const std::string &s) {
    size_t e = s.size();
    while (e > 0 && (s[e - 1] == ' ' || s[e - 1] == '\t' || s[e - 1] == '\r')) e--;
    return s.substr(0, e);
  }
  // Removes a '#' comment that is outside quotes and preceded by a space.
  static std::string strip_comment(const std::string &s) {
    char q = 0;
    for (size_t i = 0; i < s.size(); i++) {
      char c = s[i];
      if (q) {
        if (c == q) q = 0;
      } else if (c == '\'' || c == '"') {
        q = c;
      } else if (c == '#' && (i == 0 || s[i - 1] == ' ' || s[i - 1] == '\t')) {
        return s.substr(0, i);
      }
    }
    return s;
  }
  void split(const std::string &src) {
    std::istringstream in(src);
    std::string line;
    int no = 0;
    while (std::getline(in, line)) {
      no++;
      std::string t = rstrip(strip_comment(line));
      size_t ind = 0;
      while (ind < t.size() && t[ind] == ' ') ind++;
      if (ind == t.size()) continue;
      std::string body = t.substr(ind);
      if (body == "---" || body == "...") continue;
      lines_.push_back({(int)ind, body, no});
    }
  }
  bool error(int lineno, const std::string &msg) {
    if (err_.empty()) err_ = "line " + std::to_string(lineno) + ": " + msg;
    return false;
  }
  // Key of "key: ..." or "key:" outside quotes/brackets; npos if none.
  static size_t key_colon(const std::string &s) {
    char q = 0;
    int depth = 0;
    for (size_t i = 0; i < s.size(); i++) {
      char c = s[i];
      if (q) {
        if (c == q) q = 0;
        continue;
      }
      if (c == '\'' || c == '"') q = c;
      else if (c == '[' || c == '{') depth++;
      else if (c == ']' || c == '}') depth--;
      else if (c == ':' && depth == 0 && (i + 1 == s.size() || s[i + 1] == ' ')) return i;
    }
    return std::string::npos;
  }
  static std::string unquote(const std::string &s) {
    if (s.size() >= 2 && s.front() == '\'' && s.back() == '\'') {
      std::string out;
      for (size_t i = 1; i + 1 < s.size(); i++) {
        out += s[i];
        if (s[i] == '\'' && i + 2 < s.size() && s[i + 1] == '\'') i++;
      }
      return out;
    }
    if (s.size() >= 2 && s.front() == '"' && s.back() == '"') {
      std::string out;
      for (size_t i = 1; i + 1 < s.size(); i++) {
        if (s[i] == '\\' && i + 2 < s.size()) {
          char n = s[++i];
          out += n == 'n' ? '\n' : n == 't' ? '\t' : n;
        } else {
          out += s[i];
        }
      }
      return out;
    }
    return s;
  }
  // Flow value: [a, b, [c]] / {k: v} / scalar.
  bool flow(const std::string &s, size_t &i, YNode &out, int lineno) {
    while (i < s.size() && s[i] == ' ') i++;
    if (i < s.size() && s[i] == '[') {
      out.kind = YNode::kSeq;
      i++;
      for (;;) {
        while (i < s.size() && s[i] == ' ') i++;
        if (i < s.size() && s[i] == ']') {
          i++;
          return true;
        }
        auto child = std::make_unique<YNode>();
        if (!flow(s, i, *child, lineno)) return false;
        out.seq.push_back(std::move(child));
        while (i < s.size() && s[i] == ' ') i++;
        if (i < s.size() && s[i] == ',') {
          i++;
          continue;
        }
        if (i < s.size() && s[i] == ']') {
          i++;
          return true;
        }
        return error(lineno, "unterminated flow sequence");
      }
    }
    if (i < s.size() && s[i] == '{') {
      out.kind = YNode::kMap;
      i++;
      for (;;) {
        while (i < s.size() && s[i] == ' ') i++;
        if (i < s.size() && s[i] == '}') {
          i++;
          return true;
        }
        size_t k0 = i;
        while (i < s.size() && s[i] != ':' && s[i] != ',' && s[i] != '}') i++;
        if (i >= s.size() || s[i] != ':') return error(lineno, "bad flow mapping");
        std::string key = unquote(rstrip(s.substr(k0, i - k0)));
        i++;
        auto child = std::make_unique<YNode>();
        if (!flow(s, i, *child, lineno)) return false;
        out.map.emplace_back(key, std::move(child));
        while (i < s.size() && s[i] == ' ') i++;
        if (i < s.size() && s[i] == ',') {
          i++;
          continue;
        }
        if (i < s.size() && s[i] == '}') {
          i++;
          return true;
        }
        return error(lineno, "unterminated flow mapping");
      }
    }
    size_t b = i;
    if (i < s.size() && (s[i] == '\'' || s[i] == '"')) {
      char q = s[i++];
      while (i < s.size() && s[i] != q) i++;
      i = std::min(i + 1, s.size());
    } else {
      while (i < s.size() && s[i] != ',' && s[i] != ']' && s[i] != '}') i++;
    }
    out.kind = YNode::kScalar;
    out.scalar = unquote(rstrip(s.substr(b, i - b)));
    return true;
  }
  bool inline_value(const std::string &v, YNode &out, int lineno) {
    size_t i = 0;
    if (!v.empty() && (v[0] == '[' || v[0] == '{')) {
      if (!flow(v, i, out, lineno)) return false;
      while (i < v.size() && v[i] == ' ') i++;
      if (i != v.size()) return error(lineno, "trailing text after flow value");
      return true;
    }
    out.kind = YNode::kScalar;
    out.scalar = unquote(v);
    return true;
  }
  // Plain scalar continued on more-indented lines.
  void continuation(int parent_indent, std::string &scalar) {
    while (pos_ < lines_.size() && lines_[pos_].indent > parent_indent &&
           key_colon(lines_[pos_].text) == std::string::npos && lines_[pos_].text[0] != '-') {
      scalar += (scalar.empty() ? "" : " ") + lines_[pos_].text;
      pos_++;
    }
  }
  // Value of a key / item whose text after the indicator is `rest`.
  bool value(const std::string &rest, int parent_indent, YNode &out, int lineno) {
    if (!rest.empty()) {
      if (!inline_value(rest, out, lineno)) return false;
      if (out.kind == YNode::kScalar) continuation(parent_indent, out.scalar);
      return true;
    }
    if (pos_ < lines_.size() && lines_[pos_].indent > parent_indent) {
      const YLine &nx = lines_[pos_];
      if (nx.text[0] == '[' || nx.text[0] == '{') {  // flow value on its own line
        pos_++;
        return inline_value(nx.text, out, nx.lineno);
      }
      if (nx.text[0] != '-' && key_colon(nx.text) == std::string::npos) {
        out.kind = YNode::kScalar;
        out.scalar.clear();
        continuation(parent_indent, out.scalar);
        return true;
      }
      return block(nx.indent, out);
    }
    // "key:" followed by a sequence at the same indentation
    if (pos_ < lines_.size() && lines_[pos_].indent == parent_indent &&
        lines_[pos_].text.rfind("- ", 0) == 0)
      return block(parent_indent, out);
    out.kind = YNode::kNull;
    return true;
  }
  bool block(int indent, YNode &out) {
    const YLine &first = lines_[pos_];
    const bool is_seq = first.text == "-" || first.text.rfind("- ", 0) == 0;
    out.kind = is_seq ? YNode::kSeq : YNode::kMap;
    while (pos_ < lines_.size() && lines_[pos_].indent == indent) {
      const YLine ln = lines_[pos_];
      if (is_seq) {
        if (!(ln.text == "-" || ln.text.rfind("- ", 0) == 0)) break;
        pos_++;
        std::string rest = ln.text.size() > 2 ? ln.text.substr(2) : "";
        size_t lead = 0;
        while (lead < rest.size() && rest[lead] == ' ') lead++;
        rest = rest.substr(lead);
        auto child = std::make_unique<YNode>();
        if (!rest.empty() && key_colon(rest) != std::string::npos && rest[0] != '[' &&
            rest[0] != '{') {
          // "- key: value" starts a mapping indented past the dash
          int child_indent = indent + 2 + (int)lead;
          lines_.insert(lines_.begin() + (long)pos_, YLine{child_indent, rest, ln.lineno});
          if (!block(child_indent, *child)) return false;
        } else if (!value(rest, indent, *child, ln.lineno)) {
          return false;
        }
        out.seq.push_back(std::move(child));
      } else {
        size_t c = key_colon(ln.text);
        if (c == std::string::npos) return error(ln.lineno, "expected 'key: value'");
        pos_++;
        std::string key = unquote(rstrip(ln.text.substr(0, c)));
        std::string rest = c + 1 < ln.text.size() ? ln.text.substr(c + 1) : "";
        size_t lead = 0;
        while (lead < rest.size() && rest[lead] == ' ') lead++;
        rest = rest.substr(lead);
        auto child = std::make_unique<YNode>();
        if (!value(rest, indent, *child, ln.lineno)) return false;
        out.map.emplace_back(key, std::move(child));
      }
    }
    if (pos_ < lines_.size() && lines_[pos_].indent > indent)
      return error(lines_[pos_].lineno, "unexpected indentation");
    return true;
  }

  std::vector<YLine> lines_;
  size_t pos_ = 0;
  std::string err_;
};

// std::stoi / std::stof as the reference calls them (yaml_config.h:31-49):
// leading integer / float, trailing text ignored, no digits is an error.
bool parse_int(const std::string &s, long &out) {
  const char *b = s.c_str();
  char *e = nullptr;
  errno = 0;
  long v = std::strtol(b, &e, 10);
  if (e == b || errno == ERANGE || v < INT_MIN || v > INT_MAX) return false;
  out = v;
  return true;
}

bool parse_float(const std::string &s, float &out) {
  const char *b = s.c_str();
  char *e = nullptr;
  errno = 0;
  float v = std::strtof(b, &e);
  if (e == b || errno == ERANGE) return false;
  out = v;
  return true;
}

bool scalar_list(const YNode *n, std::vector<std::string> &out) {
  if (!n || n->kind != YNode::kSeq) return false;
  out.clear();
  for (auto &c : n->seq) {
    if (c->kind != YNode::kScalar) return false;
    out.push_back(c->scalar);
  }
  return true;
}

struct Rule {
  std::vector<long> relation_l, relation_r, eq, sim, conseq;
  std::vector<std::pair<long, float>> threshold;  // insertion order, first key wins
  std::string source;
  const float *find_threshold(long col) const {
    for (auto &kv : threshold)
      if (kv.first == col) return &kv.second;
    return nullptr;
  }
};

// YAML::convert<Rule>::decode, yaml_config.h:19-57.
hb_status decode_rule(const std::string &text, const std::string &name, Rule &rule) {
  YamlParser parser(text);
  std::unique_ptr<YNode> root;
  std::string err;
  if (!parser.parse(root, err)) return fail(HB_ERR_PARSE, "%s: %s", name.c_str(), err.c_str());
  if (root->size() < 2)
    return fail(HB_ERR_PARSE, "%s: invalid rule's metadata format (fewer than 2 top-level keys)",
                name.c_str());
  const YNode *pre = root->get("Preconditions");
  if (!pre || pre->kind != YNode::kMap)
    return fail(HB_ERR_PARSE, "%s: no Preconditions mapping", name.c_str());
  std::vector<std::string> rl, rr;
  const YNode *rel = pre->get("Relations");
  if (!scalar_list(rel ? rel->at(0) : nullptr, rl) || !scalar_list(rel->at(1), rr))
    return fail(HB_ERR_PARSE, "%s: Preconditions.Relations needs two lists", name.c_str());
  // yaml_config.h:30: the rule is only filled when Preconditions has exactly
  // 4 keys; an unfilled rule makes the reference read past its lists, so it
  // is rejected here.
  if (pre->size() != 4)
    return fail(HB_ERR_PLAN,
                "%s: Preconditions must have exactly 4 keys (Relations, Equalities, Sim, "
                "Threshold), found %zu",
                name.c_str(), pre->size());
  auto ints = [&](const std::vector<std::string> &v, std::vector<long> &o, const char *what) {
    o.clear();
    for (auto &s : v) {
      long x;
      if (!parse_int(s, x))
        return fail(HB_ERR_PARSE, "%s: %s: '%s' is not an integer", name.c_str(), what, s.c_str());
      o.push_back(x);
    }
    return HB_OK;
  };
  hb_status st;
  if ((st = ints(rl, rule.relation_l, "Relations[0]")) != HB_OK) return st;
  if ((st = ints(rr, rule.relation_r, "Relations[1]")) != HB_OK) return st;
  std::vector<std::string> eq, sim, thr, cons;
  const YNode *n;
  if (!(n = pre->get("Equalities")) || !scalar_list(n->at(0), eq))
    return fail(HB_ERR_PARSE, "%s: Preconditions.Equalities[0] must be a list", name.c_str());
  if (!(n = pre->get("Sim")) || !scalar_list(n->at(0), sim))
    return fail(HB_ERR_PARSE, "%s: Preconditions.Sim[0] must be a list", name.c_str());
  if (!(n = pre->get("Threshold")) || !scalar_list(n->at(0), thr))
    return fail(HB_ERR_PARSE, "%s: Preconditions.Threshold[0] must be a list", name.c_str());
  if ((st = ints(eq, rule.eq, "Equalities[0]")) != HB_OK) return st;
  if ((st = ints(sim, rule.sim, "Sim[0]")) != HB_OK) return st;
  for (size_t i = 2; i < thr.size(); i++) {
    const size_t k = 2 + (i - 2) * 2;
    if (k >= rule.sim.size())
      return fail(HB_ERR_PLAN, "%s: Threshold[0] has more entries than Sim[0] has pairs",
                  name.c_str());
    float f;
    if (!parse_float(thr[i], f))
      return fail(HB_ERR_PARSE, "%s: threshold '%s' is not a number", name.c_str(),
                  thr[i].c_str());
    if (!rule.find_threshold(rule.sim[k])) rule.threshold.emplace_back(rule.sim[k], f);
  }
  const YNode *con = root->get("Conseq");
  if (!con || !scalar_list(con->get("Equality"), cons))
    return fail(HB_ERR_PARSE, "%s: Conseq.Equality must be a list", name.c_str());
  if ((st = ints(cons, rule.conseq, "Conseq.Equality")) != HB_OK) return st;
  rule.source = name;
  return HB_OK;
}

}  // namespace

struct hb_host_plan {
  std::vector<int32_t> pred_index;
  std::vector<char> pred_type;
  std::vector<float> pred_threshold;
  int32_t n_rules = 0;
};

namespace {

// GetExecutionPlan(kEqualitiesFirst), execution_plan_generator.h:84-155.
hb_status build_plan(std::vector<Rule> &rules, hb_host_plan **out) {
  if (rules.empty()) return fail(HB_ERR_PLAN, "no rules");
  // execution_plan_generator.h:75-79: rules ordered by |Equalities[0]|,
  // descending.  Stable here so equal counts keep file-name order.
  std::stable_sort(rules.begin(), rules.end(),
                   [](const Rule &a, const Rule &b) { return a.eq.size() > b.eq.size(); });
  auto plan = std::make_unique<hb_host_plan>();
  for (const Rule &r : rules) {
    if (r.eq.size() < 2 || r.sim.size() < 2)
      return fail(HB_ERR_PLAN, "%s: Equalities[0] and Sim[0] must start with two relation ids",
                  r.source.c_str());
    if ((r.eq.size() - 2) % 2 || (r.sim.size() - 2) % 2)
      return fail(HB_ERR_PLAN, "%s: predicate lists must hold (left, right) column pairs",
                  r.source.c_str());
    if (r.eq.size() == 2 && r.sim.size() == 2)
      return fail(HB_ERR_PLAN, "%s: rule has no predicate", r.source.c_str());
    for (size_t i = 2; i < r.eq.size(); i++) {
      if (r.eq[i] < 0) return fail(HB_ERR_PLAN, "%s: negative column", r.source.c_str());
      plan->pred_index.push_back((int32_t)r.eq[i]);
      plan->pred_type.push_back(HB_EQUALITIES);
      plan->pred_threshold.push_back(1.0f);
    }
    for (size_t i = 2; i < r.sim.size(); i++) {
      if (r.sim[i] < 0) return fail(HB_ERR_PLAN, "%s: negative column", r.source.c_str());
      // The threshold map is keyed by left columns (yaml_config.h:47-50); a
      // right column missing from it would be read through end() by the
      // reference (execution_plan_generator.h:130-131).
      const float *t = r.find_threshold(r.sim[i]);
      if (!t)
        return fail(HB_ERR_PLAN,
                    "%s: column %ld has no threshold (thresholds are keyed by the left column "
                    "of each Sim pair)",
                    r.source.c_str(), r.sim[i]);
      plan->pred_index.push_back((int32_t)r.sim[i]);
      plan->pred_type.push_back(HB_SIM);
      plan->pred_threshold.push_back(*t);
    }
    plan->pred_index.push_back(HB_CHECK_POINT);
    plan->pred_type.push_back(HB_CHECK_POINT_CHAR);
    plan->pred_threshold.push_back(1.0f);
  }
  plan->n_rules = (int32_t)rules.size();
  *out = plan.release();
  return HB_OK;
}

bool read_file(const std::string &path, std::string &out) {
  std::ifstream in(path, std::ios::binary);
  if (!in) return false;
  std::ostringstream ss;
  ss << in.rdbuf();
  out = ss.str();
  return true;
}

}  // namespace

struct hb_host_table {
  std::vector<std::string> cells;  // row-major
  int64_t n_rows = 0;
  int32_t n_cols = 0;
  // column-major fixed-width copies for hb_host_table_view
  std::vector<std::vector<char>> packed;
  std::vector<const char *> bases;
  std::vector<int64_t> strides;
};

struct hb_partitions {
  std::vector<std::vector<int64_t>> l_rows, r_rows;
};

namespace {

// Cells of CSV lines, row after row, with each row's cell count.
struct CsvRows {
  std::vector<std::string> cells;
  std::vector<uint32_t> n_cells;
};

// rapidcsv parse (LabelParams(-1, -1), default SeparatorParams with the
// given separator) of buf[begin, end): quote toggling only at the start of
// a cell or inside a cell that began with a quote, CR dropped, every LF ends
// a row (quoted line breaks are off, so a quote never spans lines), "..."
// cells unquoted with "" -> ".  A row is the text up to its LF (or the end
// of the input), so ranges that start after an LF parse independently.
// Lines with no quote and no CR are split at the separator directly.
void parse_csv_range(const char *buf, size_t begin, size_t end, char sep, CsvRows &out) {
  auto unquote = [](std::string &c) {
    if (c.size() >= 2 && c.front() == '"' && c.back() == '"') {
      std::string s;
      s.reserve(c.size());
      for (size_t k = 1; k + 1 < c.size(); k++) {
        s += c[k];
        if (c[k] == '"' && k + 2 < c.size() && c[k + 1] == '"') k++;
      }
      c.swap(s);
    }
  };
  size_t i = begin;
  while (i < end) {
    const char *nl = (const char *)std::memchr(buf + i, '\n', end - i);
    const size_t le = nl ? (size_t)(nl - buf) : end;  // line [i, le)
    if (!nl) {  // the last line, without LF: a row unless it holds only CRs
      size_t k = i;
      while (k < le && buf[k] == '\r') k++;
      if (k == le) break;
    }
    const size_t before = out.cells.size();
    if (!std::memchr(buf + i, '"', le - i) && !std::memchr(buf + i, '\r', le - i)) {
      size_t s = i;
      for (;;) {
        const char *q = (const char *)std::memchr(buf + s, sep, le - s);
        const size_t ce = q ? (size_t)(q - buf) : le;
        out.cells.emplace_back(buf + s, ce - s);
        if (!q) break;
        s = ce + 1;
      }
    } else {
      std::string cell;
      bool quoted = false;
      for (size_t k = i; k < le; k++) {
        const char ch = buf[k];
        if (ch == '"') {
          if (cell.empty() || cell[0] == '"') quoted = !quoted;
          cell += ch;
        } else if (ch == sep && !quoted) {
          unquote(cell);
          out.cells.push_back(std::move(cell));
          cell.clear();
        } else if (ch != '\r') {
          cell += ch;
        }
      }
      unquote(cell);
      out.cells.push_back(std::move(cell));
    }
    out.n_cells.push_back((uint32_t)(out.cells.size() - before));
    i = le + 1;
  }
}

unsigned host_threads(size_t work, size_t per_thread) {
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  return (unsigned)std::max<size_t>(1, std::min<size_t>({hw, 16, work / per_thread}));
}

void pack_view(hb_host_table *t) {
  const int nc = t->n_cols;
  const size_t nr = (size_t)t->n_rows;
  t->packed.assign(nc, {});
  t->bases.assign(nc, nullptr);
  t->strides.assign(nc, 1);
  const unsigned nt = host_threads(nr * nc, 65536);
  std::vector<std::vector<size_t>> wt(nt, std::vector<size_t>(nc, 1));
  auto rows_of = [&](unsigned k) { return std::make_pair(nr * k / nt, nr * (k + 1) / nt); };
  auto run = [&](auto fn) {
    std::vector<std::thread> th;
    for (unsigned k = 1; k < nt; k++) th.emplace_back(fn, k);
    fn(0u);
    for (auto &x : th) x.join();
  };
  run([&](unsigned k) {  // widths: longest cell + 1 per column
    const auto rr = rows_of(k);
    for (size_t r = rr.first; r < rr.second; r++)
      for (int c = 0; c < nc; c++)
        wt[k][c] = std::max(wt[k][c], t->cells[r * nc + c].size() + 1);
  });
  for (int c = 0; c < nc; c++) {
    size_t w = 1;
    for (unsigned k = 0; k < nt; k++) w = std::max(w, wt[k][c]);
    t->packed[c].resize(std::max<size_t>(nr * w, 1));
    t->bases[c] = t->packed[c].data();
    t->strides[c] = (int64_t)w;
  }
  run([&](unsigned k) {  // zero-padded fixed-width copies
    const auto rr = rows_of(k);
    for (int c = 0; c < nc; c++) {
      const size_t w = (size_t)t->strides[c];
      char *p = t->packed[c].data();
      std::memset(p + rr.first * w, 0, (rr.second - rr.first) * w);
      for (size_t r = rr.first; r < rr.second; r++) {
        const std::string &s = t->cells[r * nc + c];
        std::memcpy(p + r * w, s.data(), s.size());
      }
    }
  });
}

}  // namespace

namespace {

// DataPartitioning (data_mngr.h:81-222) over any cell accessor
// cell(side, row, col) -> std::string_view.  std::hash<std::string_view>
// equals std::hash<std::string> on the same bytes.
template <class Cell>
hb_status partition_rows(int64_t n_l, int32_t cols_l, int64_t n_r, int32_t cols_r,
                         const hb_serialized_plan *plan, uint64_t n_partitions,
                         int64_t prefix_idx, Cell cell, hb_partitions **out) {
  if (!plan || !out) return fail(HB_ERR_INVALID_ARG, "hb_partition: null argument");
  *out = nullptr;
  if (n_partitions == 0) return fail(HB_ERR_INVALID_ARG, "n_partitions must be >= 1");
  const bool use_prefix = prefix_idx != (int64_t)INT_MAX;
  if (use_prefix && (prefix_idx < 0 || prefix_idx >= cols_l || prefix_idx >= cols_r))
    return fail(HB_ERR_INVALID_ARG, "prefix_hash_pred_idx %lld out of range",
                (long long)prefix_idx);
  // data_mngr.h:152-160: only plan slot 0 is hashed (MAX_HASH_TIMES = 1),
  // and its (left) column index is used on both sides.
  const bool hash_slot0 = plan->length > 0 && plan->pred_type[0] == HB_EQUALITIES;
  const int32_t c0 = hash_slot0 ? plan->pred_index[0] : 0;
  if (hash_slot0 && (c0 < 0 || c0 >= cols_l || c0 >= cols_r))
    return fail(HB_ERR_PLAN, "plan slot 0 column %d out of range", c0);
  const size_t P = (size_t)n_partitions;
  std::hash<std::string_view> h;
  auto bucket_of = [&](int side, int64_t row) -> size_t {
    int bucket_id = 0;
    if (hash_slot0) {
      const size_t local = h(cell(side, row, c0)) % P;
      bucket_id = (int)(((size_t)bucket_id + local) % P);
    }
    if (use_prefix) {
      const std::string_view s = cell(side, row, (int32_t)prefix_idx);
      if (s.length() > 2) {
        const size_t two_d = (h(s.substr(0, 1)) + std::hash<int>{}(bucket_id)) % P;
        bucket_id = (int)(((size_t)bucket_id + two_d) % P);
      }
    }
    return (size_t)bucket_id;
  };
  std::vector<std::vector<int64_t>> bl(P), br(P);
  for (int64_t i = 0; i < n_l; i++) bl[bucket_of(0, i)].push_back(i);
  for (int64_t i = 0; i < n_r; i++) br[bucket_of(1, i)].push_back(i);
  auto parts = std::make_unique<hb_partitions>();
  for (size_t k = 0; k < P; k++) {
    if (bl[k].empty() || br[k].empty()) continue;  // data_mngr.h:210-211
    parts->l_rows.push_back(std::move(bl[k]));
    parts->r_rows.push_back(std::move(br[k]));
  }
  *out = parts.release();
  return HB_OK;
}

}  // namespace

extern "C" {

hb_status hb_rules_parse(int32_t n_docs, const char *const *docs, hb_host_plan **out) {
  if (!out || n_docs <= 0 || !docs) return fail(HB_ERR_INVALID_ARG, "hb_rules_parse: no rules");
  *out = nullptr;
  std::vector<Rule> rules(n_docs);
  for (int i = 0; i < n_docs; i++) {
    if (!docs[i]) return fail(HB_ERR_INVALID_ARG, "hb_rules_parse: null document %d", i);
    hb_status st = decode_rule(docs[i], "rule " + std::to_string(i), rules[i]);
    if (st != HB_OK) return st;
  }
  return build_plan(rules, out);
}

hb_status hb_rules_load(const char *rule_dir, hb_host_plan **out) {
  if (!out || !rule_dir) return fail(HB_ERR_INVALID_ARG, "hb_rules_load: null argument");
  *out = nullptr;
  DIR *d = opendir(rule_dir);
  if (!d) return fail(HB_ERR_IO, "cannot open rule directory '%s': %s", rule_dir,
                      std::strerror(errno));
  std::vector<std::string> files;
  while (dirent *e = readdir(d)) {
    std::string name = e->d_name;
    if (name == "." || name == "..") continue;
    std::string path = std::string(rule_dir) + "/" + name;
    struct stat sb;
    if (stat(path.c_str(), &sb) == 0 && S_ISREG(sb.st_mode)) files.push_back(path);
  }
  closedir(d);
  // directory_iterator order is unspecified (execution_plan_generator.h:66);
  // file names are sorted so plans are reproducible.
  std::sort(files.begin(), files.end());
  if (files.empty()) return fail(HB_ERR_PLAN, "rule directory '%s' has no files", rule_dir);
  std::vector<Rule> rules(files.size());
  for (size_t i = 0; i < files.size(); i++) {
    std::string text;
    if (!read_file(files[i], text)) return fail(HB_ERR_IO, "cannot read '%s'", files[i].c_str());
    hb_status st = decode_rule(text, files[i], rules[i]);
    if (st != HB_OK) return st;
  }
  return build_plan(rules, out);
}

hb_status hb_host_plan_view(const hb_host_plan *p, hb_serialized_plan *v) {
  if (!p || !v) return fail(HB_ERR_INVALID_ARG, "hb_host_plan_view: null argument");
  v->n_rules = p->n_rules;
  v->length = (int32_t)p->pred_index.size();
  v->pred_index = p->pred_index.data();
  v->pred_type = p->pred_type.data();
  v->pred_threshold = p->pred_threshold.data();
  return HB_OK;
}

void hb_host_plan_free(hb_host_plan *p) { delete p; }

hb_status hb_csv_load(const char *path, char sep, hb_host_table **out) {
  if (!path || !out) return fail(HB_ERR_INVALID_ARG, "hb_csv_load: null argument");
  *out = nullptr;
  std::string buf;
  {
    FILE *f = std::fopen(path, "rb");
    if (!f) return fail(HB_ERR_IO, "cannot read '%s'", path);
    struct stat sb;
    if (fstat(fileno(f), &sb) == 0 && sb.st_size > 0) buf.resize((size_t)sb.st_size);
    const size_t got = buf.empty() ? 0 : std::fread(&buf[0], 1, buf.size(), f);
    buf.resize(got);
    char tmp[65536];  // whatever the size did not cover (a file still growing, a pipe)
    for (size_t n; (n = std::fread(tmp, 1, sizeof tmp, f)) > 0;) buf.append(tmp, n);
    const bool err = std::ferror(f) != 0;
    std::fclose(f);
    if (err) return fail(HB_ERR_IO, "cannot read '%s'", path);
  }
  size_t begin = 0;
  if (buf.size() >= 3 && (unsigned char)buf[0] == 0xEF && (unsigned char)buf[1] == 0xBB &&
      (unsigned char)buf[2] == 0xBF)
    begin = 3;
  // chunks that start after an LF parse independently (parse_csv_range)
  const unsigned nt = host_threads(buf.size() - begin, 1u << 20);
  std::vector<size_t> cut(nt + 1, buf.size());
  cut[0] = begin;
  for (unsigned k = 1; k < nt; k++) {
    size_t c = std::max(cut[k - 1], begin + (buf.size() - begin) * k / nt);
    const void *nl = c < buf.size() ? std::memchr(buf.data() + c, '\n', buf.size() - c) : nullptr;
    cut[k] = nl ? (size_t)((const char *)nl - buf.data()) + 1 : buf.size();
  }
  std::vector<CsvRows> part(nt);
  {
    std::vector<std::thread> th;
    for (unsigned k = 1; k < nt; k++)
      th.emplace_back([&, k] { parse_csv_range(buf.data(), cut[k], cut[k + 1], sep, part[k]); });
    parse_csv_range(buf.data(), cut[0], cut[1], sep, part[0]);
    for (auto &x : th) x.join();
  }
  auto t = std::make_unique<hb_host_table>();
  size_t n_rows = 0;
  for (const auto &p : part) n_rows += p.n_cells.size();
  uint32_t n_cols = 0;
  for (const auto &p : part)
    if (!p.n_cells.empty()) {
      n_cols = p.n_cells[0];
      break;
    }
  if (n_rows == 0 || n_cols == 0) return fail(HB_ERR_PARSE, "'%s' holds no rows", path);
  size_t row0 = 0;  // rows before each chunk
  for (const auto &p : part) {
    for (size_t r = 0; r < p.n_cells.size(); r++)
      if (p.n_cells[r] < n_cols)
        return fail(HB_ERR_PARSE, "'%s' line %zu has %u cells, the first line has %u", path,
                    row0 + r + 1, p.n_cells[r], n_cols);
    row0 += p.n_cells.size();
  }
  t->n_rows = (int64_t)n_rows;
  t->n_cols = (int32_t)n_cols;
  t->cells.resize(n_rows * n_cols);
  {  // each chunk's first n_cols cells of every row into place (extra cells are dropped)
    std::vector<std::thread> th;
    auto place = [&](unsigned k, size_t r0) {
      CsvRows &p = part[k];
      size_t src = 0;
      for (size_t r = 0; r < p.n_cells.size(); r++) {
        for (uint32_t c = 0; c < n_cols; c++)
          t->cells[(r0 + r) * n_cols + c] = std::move(p.cells[src + c]);
        src += p.n_cells[r];
      }
      CsvRows().cells.swap(p.cells);
    };
    size_t r0 = 0;
    for (unsigned k = 0; k < nt; k++) {
      if (k) th.emplace_back(place, k, r0);
      r0 += part[k].n_cells.size();
    }
    place(0, 0);
    for (auto &x : th) x.join();
  }
  pack_view(t.get());
  *out = t.release();
  return HB_OK;
}

hb_status hb_host_table_view(const hb_host_table *t, hb_table_view *v) {
  if (!t || !v) return fail(HB_ERR_INVALID_ARG, "hb_host_table_view: null argument");
  v->n_rows = t->n_rows;
  v->n_cols = t->n_cols;
  v->col_base = t->bases.data();
  v->col_stride = t->strides.data();
  return HB_OK;
}

const char *hb_host_table_cell(const hb_host_table *t, int64_t row, int32_t col) {
  if (!t || row < 0 || row >= t->n_rows || col < 0 || col >= t->n_cols) return nullptr;
  return t->cells[(size_t)row * t->n_cols + col].c_str();
}

void hb_host_table_free(hb_host_table *t) { delete t; }


hb_status hb_partition(const hb_host_table *l, const hb_host_table *r,
                       const hb_serialized_plan *plan, uint64_t n_partitions, int64_t prefix_idx,
                       hb_partitions **out) {
  if (!l || !r) return fail(HB_ERR_INVALID_ARG, "hb_partition: null argument");
  auto cell = [&](int side, int64_t row, int32_t col) -> std::string_view {
    const hb_host_table *t = side ? r : l;
    return t->cells[(size_t)row * t->n_cols + col];
  };
  return partition_rows(l->n_rows, l->n_cols, r->n_rows, r->n_cols, plan, n_partitions,
                        prefix_idx, cell, out);
}

hb_status hb_partition_views(const hb_table_view *l, const hb_table_view *r,
                             const hb_serialized_plan *plan, uint64_t n_partitions,
                             int64_t prefix_idx, hb_partitions **out) {
  if (!l || !r) return fail(HB_ERR_INVALID_ARG, "hb_partition_views: null argument");
  for (const hb_table_view *v : {l, r})
    if (v->n_rows < 0 || v->n_cols <= 0 || !v->col_base || !v->col_stride)
      return fail(HB_ERR_INVALID_ARG, "hb_partition_views: bad table view");
  auto cell = [&](int side, int64_t row, int32_t col) -> std::string_view {
    const hb_table_view *v = side ? r : l;
    const char *s = v->col_base[col] + row * v->col_stride[col];
    return std::string_view(s, strnlen(s, (size_t)v->col_stride[col]));
  };
  return partition_rows(l->n_rows, l->n_cols, r->n_rows, r->n_cols, plan, n_partitions,
                        prefix_idx, cell, out);
}

int64_t hb_partitions_count(const hb_partitions *p) { return p ? (int64_t)p->l_rows.size() : 0; }

hb_status hb_partitions_get(const hb_partitions *p, int64_t k, const int64_t **l_rows,
                            int64_t *n_l, const int64_t **r_rows, int64_t *n_r) {
  if (!p || k < 0 || k >= (int64_t)p->l_rows.size())
    return fail(HB_ERR_INVALID_ARG, "partition %lld out of range", (long long)k);
  if (l_rows) *l_rows = p->l_rows[k].data();
  if (n_l) *n_l = (int64_t)p->l_rows[k].size();
  if (r_rows) *r_rows = p->r_rows[k].data();
  if (n_r) *n_r = (int64_t)p->r_rows[k].size();
  return HB_OK;
}

void hb_partitions_free(hb_partitions *p) { delete p; }

hb_status hb_write_pairs(const char *path, int32_t append, const hb_host_table *l,
                         const hb_host_table *r, const int64_t *l_rows, const int64_t *r_rows,
                         const uint32_t *pairs, uint64_t n_pairs) {
  if (!path || !l || !r || (n_pairs && !pairs))
    return fail(HB_ERR_INVALID_ARG, "hb_write_pairs: null argument");
  FILE *f = std::fopen(path, append ? "ab" : "wb");
  if (!f) return fail(HB_ERR_IO, "cannot open '%s': %s", path, std::strerror(errno));
  std::string buf;
  buf.reserve(1 << 20);
  for (uint64_t i = 0; i < n_pairs; i++) {
    const int64_t lr = l_rows ? l_rows[pairs[2 * i]] : (int64_t)pairs[2 * i];
    const int64_t rr = r_rows ? r_rows[pairs[2 * i + 1]] : (int64_t)pairs[2 * i + 1];
    if (lr < 0 || lr >= l->n_rows || rr < 0 || rr >= r->n_rows) {
      std::fclose(f);
      return fail(HB_ERR_INVALID_ARG, "pair %llu refers to a missing row",
                  (unsigned long long)i);
    }
    buf += l->cells[(size_t)lr * l->n_cols];
    buf += ',';
    buf += r->cells[(size_t)rr * r->n_cols];
    buf += '\n';
    if (buf.size() > (1 << 20)) {
      std::fwrite(buf.data(), 1, buf.size(), f);
      buf.clear();
    }
  }
  std::fwrite(buf.data(), 1, buf.size(), f);
  if (std::fclose(f) != 0) return fail(HB_ERR_IO, "write to '%s' failed", path);
  return HB_OK;
}

}  // extern "C"
