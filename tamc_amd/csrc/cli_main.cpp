// cli_main.cpp -- the `tamc` command line, a drop-in for the reference's binary (src/main.rs,
// src/lib.rs:79-132): the same arguments, YAML method files, instance files and output files, with
// the PT-ICM and SA solves running on the GPU through libtamc_gpu.
//
//   tamc [--qubo] [--suscepts FILE...] [--sample-output FILE] <method-file> <instance-file> <output-file>
//
// PT (src/pt.rs:609-655): output-file gets PtIcmMinResults as YAML; the sample file (default
// samples.bin) gets PtIcmThermalSamples as bincode, or as a pickle when its name ends in .pkl.
// SA (src/sa.rs:216-270): output-file gets AnnealMinResults when it ends in .yml or .pkl; the sample
// file gets the AnnealState of the ground states.
// GLA (src/gla.rs:122-196): output-file gets GlaResults, bincode or (for .pkl) pickle; the
// partitions' sub-instances are solved in one batched launch.
//
// The reference seeds its generators from the OS; so does this binary, unless TAMC_SEED is set.
// --dry-run (not in the reference) parses everything and prints the resolved run as JSON without
// touching the GPU.
#include <errno.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <chrono>
#include <fstream>
#include <memory>
#include <random>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/tamc_gpu.h"

namespace {

struct Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// ------------------------------------------------------------------ minimal YAML reader ----
// Enough of YAML 1.1 for method files: block mappings and sequences by indentation, flow
// sequences/mappings ([..], {..}), plain and quoted scalars, comments, '---' / '...' markers.
struct Node {
  enum Kind { Null, Scalar, Map, Seq } kind = Null;
  std::string s;
  bool quoted = false;
  std::vector<std::pair<std::string, Node>> map;
  std::vector<Node> seq;
  const Node *get(const std::string &k) const {
    for (auto &kv : map)
      if (kv.first == k) return &kv.second;
    return nullptr;
  }
};

struct Line {
  int indent;
  std::string text;
  int number;
};

std::string trim(const std::string &s) {
  size_t a = 0, b = s.size();
  while (a < b && (s[a] == ' ' || s[a] == '\t' || s[a] == '\r')) ++a;
  while (b > a && (s[b - 1] == ' ' || s[b - 1] == '\t' || s[b - 1] == '\r')) --b;
  return s.substr(a, b - a);
}

std::string strip_comment(const std::string &s) {
  char q = 0;
  for (size_t k = 0; k < s.size(); ++k) {
    const char c = s[k];
    if (q) {
      if (c == q) q = 0;
    } else if (c == '"' || c == '\'') {
      q = c;
    } else if (c == '#' && (k == 0 || s[k - 1] == ' ' || s[k - 1] == '\t')) {
      return s.substr(0, k);
    }
  }
  return s;
}

// position of the ':' that separates a mapping key, or npos
size_t key_colon(const std::string &t) {
  char q = 0;
  int depth = 0;
  for (size_t k = 0; k < t.size(); ++k) {
    const char c = t[k];
    if (q) {
      if (c == q) q = 0;
      continue;
    }
    if (c == '"' || c == '\'') q = c;
    else if (c == '[' || c == '{') ++depth;
    else if (c == ']' || c == '}') --depth;
    else if (c == ':' && depth == 0 && (k + 1 == t.size() || t[k + 1] == ' ')) return k;
  }
  return std::string::npos;
}

Node scalar(const std::string &raw) {
  Node n;
  const std::string t = trim(raw);
  if (t.empty() || t == "~" || t == "null" || t == "Null" || t == "NULL") return n;
  n.kind = Node::Scalar;
  if (t.size() >= 2 && (t[0] == '"' || t[0] == '\'') && t.back() == t[0]) {
    n.s = t.substr(1, t.size() - 2);
    n.quoted = true;
  } else {
    n.s = t;
  }
  return n;
}

struct Flow {
  const std::string &t;
  size_t k = 0;
  int line;
  void ws() {
    while (k < t.size() && (t[k] == ' ' || t[k] == '\t')) ++k;
  }
  Node value() {
    ws();
    if (k < t.size() && t[k] == '[') {
      ++k;
      Node n;
      n.kind = Node::Seq;
      ws();
      if (k < t.size() && t[k] == ']') {
        ++k;
        return n;
      }
      for (;;) {
        n.seq.push_back(value());
        ws();
        if (k < t.size() && t[k] == ',') {
          ++k;
          continue;
        }
        if (k < t.size() && t[k] == ']') {
          ++k;
          return n;
        }
        throw Error("line " + std::to_string(line) + ": unterminated flow sequence");
      }
    }
    if (k < t.size() && t[k] == '{') {
      ++k;
      Node n;
      n.kind = Node::Map;
      ws();
      if (k < t.size() && t[k] == '}') {
        ++k;
        return n;
      }
      for (;;) {
        ws();
        const size_t a = k;
        while (k < t.size() && t[k] != ':') ++k;
        if (k >= t.size()) throw Error("line " + std::to_string(line) + ": expected ':' in flow mapping");
        const std::string key = scalar(t.substr(a, k - a)).s;
        ++k;
        n.map.push_back({key, value()});
        ws();
        if (k < t.size() && t[k] == ',') {
          ++k;
          continue;
        }
        if (k < t.size() && t[k] == '}') {
          ++k;
          return n;
        }
        throw Error("line " + std::to_string(line) + ": unterminated flow mapping");
      }
    }
    const size_t a = k;
    char q = 0;
    while (k < t.size()) {
      const char c = t[k];
      if (q) {
        if (c == q) q = 0;
      } else if (c == '"' || c == '\'') {
        q = c;
      } else if (c == ',' || c == ']' || c == '}') {
        break;
      }
      ++k;
    }
    return scalar(t.substr(a, k - a));
  }
};

Node inline_value(const std::string &text, int line) {
  const std::string t = trim(text);
  if (!t.empty() && (t[0] == '[' || t[0] == '{')) {
    Flow f{t, 0, line};
    Node n = f.value();
    f.ws();
    if (f.k != t.size()) throw Error("line " + std::to_string(line) + ": trailing characters after flow value");
    return n;
  }
  return scalar(t);
}

struct Yaml {
  std::vector<Line> lines;
  size_t i = 0;

  explicit Yaml(const std::string &src) {
    std::istringstream in(src);
    std::string raw;
    int no = 0;
    while (std::getline(in, raw)) {
      ++no;
      const std::string s = strip_comment(raw);
      const std::string t = trim(s);
      if (t.empty() || t == "---" || t == "...") continue;
      if (t.rfind("--- ", 0) == 0) {  // "--- value" on the marker line
        lines.push_back({0, trim(t.substr(4)), no});
        continue;
      }
      int ind = 0;
      while (ind < (int)s.size() && s[ind] == ' ') ++ind;
      lines.push_back({ind, t, no});
    }
  }

  static bool is_item(const std::string &t) { return t == "-" || t.rfind("- ", 0) == 0; }

  Node block(int indent) {
    if (i >= lines.size()) return Node();
    if (is_item(lines[i].text)) return sequence(lines[i].indent);
    if (key_colon(lines[i].text) != std::string::npos) return mapping(lines[i].indent);
    Node n = inline_value(lines[i].text, lines[i].number);
    ++i;
    (void)indent;
    return n;
  }

  Node sequence(int indent) {
    Node n;
    n.kind = Node::Seq;
    while (i < lines.size() && lines[i].indent == indent && is_item(lines[i].text)) {
      const std::string rest = lines[i].text.size() > 1 ? trim(lines[i].text.substr(1)) : "";
      if (rest.empty()) {
        ++i;
        n.seq.push_back(i < lines.size() && lines[i].indent > indent ? block(lines[i].indent) : Node());
      } else {
        // the item's content continues as if it started on its own line, two columns in
        lines[i].indent = indent + 2;
        lines[i].text = rest;
        n.seq.push_back(block(indent + 2));
      }
    }
    return n;
  }

  Node mapping(int indent) {
    Node n;
    n.kind = Node::Map;
    while (i < lines.size() && lines[i].indent == indent && !is_item(lines[i].text)) {
      const std::string &t = lines[i].text;
      const size_t c = key_colon(t);
      if (c == std::string::npos)
        throw Error("line " + std::to_string(lines[i].number) + ": expected 'key: value'");
      const std::string key = scalar(t.substr(0, c)).s;
      const std::string val = c + 1 < t.size() ? trim(t.substr(c + 1)) : "";
      const int no = lines[i].number;
      ++i;
      if (!val.empty()) {
        n.map.push_back({key, inline_value(val, no)});
      } else if (i < lines.size() && (lines[i].indent > indent || (lines[i].indent == indent && is_item(lines[i].text)))) {
        n.map.push_back({key, block(lines[i].indent)});
      } else {
        n.map.push_back({key, Node()});
      }
    }
    return n;
  }

  Node parse() {
    if (lines.empty()) return Node();
    Node n = block(lines[0].indent);
    if (i != lines.size()) throw Error("line " + std::to_string(lines[i].number) + ": unexpected indentation");
    return n;
  }
};

// ------------------------------------------------------------------------ method structs ----
struct BetaOpt {
  bool geometric = true;
  float beta_min = 0, beta_max = 0;
  uint32_t num_beta = 0;
  std::vector<float> arr;
};

template <class T>
struct Opt {
  bool has = false;
  T v{};
};

struct PtMethod {
  uint32_t num_sweeps = 0;
  double warmup_fraction = 0;
  BetaOpt beta;
  Opt<float> lo_beta;
  Opt<uint32_t> lo_num_beta;
  bool icm = false;
  uint32_t num_replica_chains = 0, threads = 0;
  Opt<uint32_t> sample, sample_states;
  Opt<uint32_t> sample_limiting;
};

struct SaMethod {
  BetaOpt beta;
  uint32_t num_replicas = 0, threads = 0;
};

// GlaParams (src/gla.rs:28-33)
struct GlaMethod {
  uint32_t num_threads = 0;
  PtMethod pt;
  std::string partition_file;
};

const Node &need(const Node &m, const char *k, const char *ctx) {
  const Node *v = m.get(k);
  if (!v) throw Error(std::string(ctx) + ": missing field `" + k + "`");
  return *v;
}

uint64_t as_uint(const Node &n, const char *k, uint64_t maxv) {
  if (n.kind != Node::Scalar) throw Error(std::string("invalid type for `") + k + "`, expected an integer");
  char *e = nullptr;
  errno = 0;
  const unsigned long long v = strtoull(n.s.c_str(), &e, 10);
  if (n.s.empty() || *e || n.s[0] == '-' || errno || v > maxv)
    throw Error(std::string("invalid value for `") + k + "`: " + n.s);
  return v;
}

double as_float(const Node &n, const char *k) {
  if (n.kind != Node::Scalar) throw Error(std::string("invalid type for `") + k + "`, expected a number");
  std::string s = n.s;
  if (s == ".inf" || s == ".Inf" || s == "+.inf") return INFINITY;
  if (s == "-.inf" || s == "-.Inf") return -INFINITY;
  if (s == ".nan" || s == ".NaN") return NAN;
  char *e = nullptr;
  const double v = strtod(s.c_str(), &e);
  if (s.empty() || *e) throw Error(std::string("invalid value for `") + k + "`: " + n.s);
  return v;
}

bool as_bool(const Node &n, const char *k) {
  if (n.kind == Node::Scalar && (n.s == "true" || n.s == "True" || n.s == "TRUE")) return true;
  if (n.kind == Node::Scalar && (n.s == "false" || n.s == "False" || n.s == "FALSE")) return false;
  throw Error(std::string("invalid type for `") + k + "`, expected a boolean");
}

template <class T>
Opt<T> opt_uint(const Node &m, const char *k, uint64_t maxv) {
  Opt<T> o;
  const Node *v = m.get(k);
  if (v && v->kind != Node::Null) {
    o.has = true;
    o.v = (T)as_uint(*v, k, maxv);
  }
  return o;
}

BetaOpt parse_beta(const Node &n) {
  BetaOpt b;
  if (n.kind != Node::Map || n.map.size() != 1) throw Error("`beta`: expected `Geometric` or `Arr`");
  const std::string &tag = n.map[0].first;
  const Node &v = n.map[0].second;
  if (tag == "Geometric") {
    if (v.kind != Node::Map) throw Error("`Geometric`: expected beta_min, beta_max, num_beta");
    b.geometric = true;
    b.beta_min = (float)as_float(need(v, "beta_min", "Geometric"), "beta_min");
    b.beta_max = (float)as_float(need(v, "beta_max", "Geometric"), "beta_max");
    b.num_beta = (uint32_t)as_uint(need(v, "num_beta", "Geometric"), "num_beta", 0xFFFFFFFFull);
  } else if (tag == "Arr") {
    if (v.kind != Node::Seq) throw Error("`Arr`: expected a sequence of numbers");
    b.geometric = false;
    for (auto &x : v.seq) b.arr.push_back((float)as_float(x, "Arr"));
  } else {
    throw Error("`beta`: unknown variant `" + tag + "`, expected `Geometric` or `Arr`");
  }
  return b;
}

PtMethod parse_pt(const Node &m) {
  if (m.kind != Node::Map) throw Error("PT: expected a mapping");
  PtMethod p;
  p.num_sweeps = (uint32_t)as_uint(need(m, "num_sweeps", "PT"), "num_sweeps", 0xFFFFFFFFull);
  p.warmup_fraction = as_float(need(m, "warmup_fraction", "PT"), "warmup_fraction");
  p.beta = parse_beta(need(m, "beta", "PT"));
  const Node *lb = m.get("lo_beta");
  if (lb && lb->kind != Node::Null) {
    p.lo_beta.has = true;
    p.lo_beta.v = (float)as_float(*lb, "lo_beta");
  }
  p.lo_num_beta = opt_uint<uint32_t>(m, "lo_num_beta", 0xFFFFFFFFull);
  p.icm = as_bool(need(m, "icm", "PT"), "icm");
  p.num_replica_chains = (uint32_t)as_uint(need(m, "num_replica_chains", "PT"), "num_replica_chains", 0xFFFFFFFFull);
  p.threads = (uint32_t)as_uint(need(m, "threads", "PT"), "threads", 0xFFFFFFFFull);
  p.sample = opt_uint<uint32_t>(m, "sample", 0xFFFFFFFFull);
  p.sample_states = opt_uint<uint32_t>(m, "sample_states", 0xFFFFFFFFull);
  p.sample_limiting = opt_uint<uint32_t>(m, "sample_limiting", 255);
  return p;
}

GlaMethod parse_gla(const Node &m);

SaMethod parse_sa(const Node &m) {
  if (m.kind != Node::Map) throw Error("SA: expected a mapping");
  SaMethod s;
  s.beta = parse_beta(need(m, "beta", "SA"));
  s.num_replicas = (uint32_t)as_uint(need(m, "num_replicas", "SA"), "num_replicas", 0xFFFFFFFFull);
  s.threads = (uint32_t)as_uint(need(m, "threads", "SA"), "threads", 0xFFFFFFFFull);
  return s;
}

GlaMethod parse_gla(const Node &m) {
  if (m.kind != Node::Map) throw Error("GLA: expected a mapping");
  GlaMethod g;
  g.num_threads = (uint32_t)as_uint(need(m, "num_threads", "GLA"), "num_threads", 0xFFFFFFFFull);
  g.pt = parse_pt(need(m, "pt_params", "GLA"));
  const Node &f = need(m, "partition_file", "GLA");
  if (f.kind != Node::Scalar) throw Error("invalid type for `partition_file`, expected a string");
  g.partition_file = f.s;
  return g;
}

// ------------------------------------------------------------------------- number text ----
// Shortest decimal digits that read back to the same float.
std::string shortest(double v, bool is_f32) {
  char buf[64];
  for (int p = 1; p <= 17; ++p) {
    snprintf(buf, sizeof buf, "%.*e", p - 1, v);
    if (is_f32 ? ((float)strtod(buf, nullptr) == (float)v) : (strtod(buf, nullptr) == v)) break;
  }
  return buf;  // d.ddde[+-]xx
}

// digits/exponent of a "%e" string -> plain decimal text
std::string plain(const std::string &e, bool keep_point_zero) {
  const bool neg = e[0] == '-';
  const std::string m = e.substr(neg ? 1 : 0, e.find('e') - (neg ? 1 : 0));
  const int ex = atoi(e.c_str() + e.find('e') + 1);
  std::string digits;
  for (char c : m)
    if (c != '.') digits += c;
  while (digits.size() > 1 && digits.back() == '0') digits.pop_back();
  const int point = ex + 1;  // digits before the decimal point
  std::string out;
  if (point <= 0) {
    out = "0." + std::string(-point, '0') + digits;
  } else if (point >= (int)digits.size()) {
    out = digits + std::string(point - digits.size(), '0');
    if (keep_point_zero) out += ".0";
  } else {
    out = digits.substr(0, point) + "." + digits.substr(point);
  }
  return (neg ? "-" : "") + out;
}

// Rust's `{}` for floats: shortest round-trip digits, never an exponent, no ".0" for integers.
std::string rust_display(double v, bool is_f32) {
  if (v != v) return "NaN";
  if (isinf(v)) return v > 0 ? "inf" : "-inf";
  if (v == 0) return signbit(v) ? "-0" : "0";
  return plain(shortest(v, is_f32), false);
}

// YAML float text: shortest round-trip digits, ".0" kept on integers, exponent form outside
// [1e-5, 1e16).
std::string yaml_float(double v, bool is_f32) {
  if (v != v) return ".nan";
  if (isinf(v)) return v > 0 ? ".inf" : "-.inf";
  if (v == 0) return signbit(v) ? "-0.0" : "0.0";
  const double a = fabs(v);
  const std::string e = shortest(v, is_f32);
  if (a >= 1e-5 && a < 1e16) return plain(e, true);
  const bool neg = e[0] == '-';
  std::string m = e.substr(neg ? 1 : 0, e.find('e') - (neg ? 1 : 0));
  if (m.find('.') != std::string::npos) {
    while (m.back() == '0') m.pop_back();
    if (m.back() == '.') m.pop_back();
  }
  return (neg ? "-" : "") + m + "e" + std::to_string(atoi(e.c_str() + e.find('e') + 1));
}

// -------------------------------------------------------------------------- YAML writer ----
struct YamlOut {
  std::string s = "---\n";
  void line(int ind, const std::string &t) { s += std::string(ind, ' ') + t + "\n"; }
};

void yaml_beta(YamlOut &y, int ind, const BetaOpt &b) {
  y.line(ind, "beta:");
  if (b.geometric) {
    y.line(ind + 2, "Geometric:");
    y.line(ind + 4, "beta_min: " + yaml_float(b.beta_min, true));
    y.line(ind + 4, "beta_max: " + yaml_float(b.beta_max, true));
    y.line(ind + 4, "num_beta: " + std::to_string(b.num_beta));
  } else if (b.arr.empty()) {
    y.line(ind + 2, "Arr: []");
  } else {
    y.line(ind + 2, "Arr:");
    for (float x : b.arr) y.line(ind + 4, "- " + yaml_float(x, true));
  }
}

template <class T, class F>
void yaml_list(YamlOut &y, int ind, const char *key, const T *v, size_t n, F fmt) {
  if (n == 0) {
    y.line(ind, std::string(key) + ": []");
    return;
  }
  y.line(ind, std::string(key) + ":");
  for (size_t k = 0; k < n; ++k) y.line(ind + 2, "- " + fmt(v[k]));
}

std::string opt_text(const Opt<uint32_t> &o) { return o.has ? std::to_string(o.v) : "~"; }

// PtIcmMinResults (src/pt.rs:96-102, MinResults flattened from src/ising_results.rs:4-13)
// PtIcmParams (src/pt.rs:259-272) at indentation ind
void pt_params_yaml(YamlOut &y, int ind, const PtMethod &p) {
  y.line(ind, "num_sweeps: " + std::to_string(p.num_sweeps));
  y.line(ind, "warmup_fraction: " + yaml_float(p.warmup_fraction, false));
  yaml_beta(y, ind, p.beta);
  y.line(ind, "lo_beta: " + (p.lo_beta.has ? yaml_float(p.lo_beta.v, true) : std::string("~")));
  y.line(ind, "lo_num_beta: " + opt_text(p.lo_num_beta));
  y.line(ind, std::string("icm: ") + (p.icm ? "true" : "false"));
  y.line(ind, "num_replica_chains: " + std::to_string(p.num_replica_chains));
  y.line(ind, "threads: " + std::to_string(p.threads));
  y.line(ind, "sample: " + opt_text(p.sample));
  y.line(ind, "sample_states: " + opt_text(p.sample_states));
  y.line(ind, "sample_limiting: " + opt_text(p.sample_limiting));
}

std::string pt_min_results_yaml(const PtMethod &p, const tamc_pt_result &r) {
  YamlOut y;
  y.line(0, "params:");
  pt_params_yaml(y, 2, p);
  y.line(0, "timing: " + yaml_float(r.timing_us, false));
  yaml_list(y, 0, "gs_time_steps", r.gs_steps, r.n_gs, [](uint32_t v) { return std::to_string(v); });
  yaml_list(y, 0, "gs_energies", r.gs_energies, r.n_gs, [](float v) { return yaml_float(v, true); });
  if (r.n_gs == 0) {
    y.line(0, "gs_states: []");
  } else {
    y.line(0, "gs_states:");
    for (uint32_t k = 0; k < r.n_gs; ++k) {
      if (r.gs_words == 0) {
        y.line(2, "- []");
        continue;
      }
      for (uint32_t w = 0; w < r.gs_words; ++w)
        y.line(w == 0 ? 2 : 4, std::string(w == 0 ? "- - " : "- ") + std::to_string(r.gs_states[(size_t)k * r.gs_words + w]));
    }
  }
  y.line(0, "num_measurements: 0");
  y.line(0, "instance_size: " + std::to_string(r.n));
  yaml_list(y, 0, "acceptance_counts", r.acceptance_counts, r.B, [](uint32_t v) { return std::to_string(v); });
  return y.s;
}

// AnnealMinResults (src/sa.rs:28-34)
std::string sa_min_results_yaml(const SaMethod &m, const tamc_sa_result &r, double timing) {
  YamlOut y;
  y.line(0, "params:");
  yaml_beta(y, 2, m.beta);
  y.line(2, "num_replicas: " + std::to_string(m.num_replicas));
  y.line(2, "threads: " + std::to_string(m.threads));
  y.line(0, "timing: " + yaml_float(timing, false));
  y.line(0, "min_energy: " + yaml_float(r.min_energy, true));
  yaml_list(y, 0, "energies", r.energies, r.R, [](float v) { return yaml_float(v, true); });
  return y.s;
}

// AnnealMinResults as a pickle (protocol 3; structs as dicts, the beta enum as {variant: value})
std::string sa_min_results_pickle(const SaMethod &m, const tamc_sa_result &r, double timing) {
  std::string b;
  auto op = [&](int c) { b.push_back((char)c); };
  auto str = [&](const std::string &s) {
    op('X');
    const uint32_t n = (uint32_t)s.size();
    b.append((const char *)&n, 4);
    b += s;
  };
  auto integer = [&](uint64_t v) {
    if (v < 256) {
      op('K');
      op((int)v);
    } else if (v < 65536) {
      op('M');
      op((int)(v & 0xFF));
      op((int)(v >> 8));
    } else if (v <= 0x7FFFFFFFull) {
      op('J');
      const int32_t w = (int32_t)v;
      b.append((const char *)&w, 4);
    } else {
      op(0x8a);
      op(8);
      b.append((const char *)&v, 8);
    }
  };
  auto real = [&](double v) {
    op('G');
    uint64_t u;
    memcpy(&u, &v, 8);
    for (int k = 7; k >= 0; --k) op((int)((u >> (8 * k)) & 0xFF));
  };
  op(0x80);
  op(3);
  op('}');
  op('(');
  str("params");
  op('}');
  op('(');
  str("beta");
  op('}');
  op('(');
  if (m.beta.geometric) {
    str("Geometric");
    op('}');
    op('(');
    str("beta_min");
    real(m.beta.beta_min);
    str("beta_max");
    real(m.beta.beta_max);
    str("num_beta");
    integer(m.beta.num_beta);
    op('u');
  } else {
    str("Arr");
    op(']');
    if (!m.beta.arr.empty()) {
      op('(');
      for (float x : m.beta.arr) real(x);
      op('e');
    }
  }
  op('u');
  str("num_replicas");
  integer(m.num_replicas);
  str("threads");
  integer(m.threads);
  op('u');
  str("timing");
  real(timing);
  str("min_energy");
  real(r.min_energy);
  str("energies");
  op(']');
  if (r.R) {
    op('(');
    for (uint32_t k = 0; k < r.R; ++k) real(r.energies[k]);
    op('e');
  }
  op('u');
  op('.');
  return b;
}

void write_text(const std::string &path, const std::string &s, const char *what) {
  std::ofstream f(path, std::ios::binary);
  if (!f) throw Error(std::string("Failed to create ") + what + " " + path);
  f << s;
  if (!f) throw Error(std::string("Failed to write ") + what + " " + path);
}

bool ends_with(const std::string &s, const char *ext) {
  // Path::extension: the part after the last '.' of the file name
  const size_t slash = s.find_last_of('/');
  const std::string name = slash == std::string::npos ? s : s.substr(slash + 1);
  const size_t dot = name.find_last_of('.');
  return dot != std::string::npos && dot > 0 && name.substr(dot + 1) == ext;
}

void check(tamc_status st) {
  if (st != TAMC_OK) throw Error(tamc_last_error());
}

uint64_t run_seed() {
  const char *s = getenv("TAMC_SEED");
  if (s && *s) return strtoull(s, nullptr, 0);
  std::random_device rd;
  return ((uint64_t)rd() << 32) ^ rd() ^ (uint64_t)std::chrono::steady_clock::now().time_since_epoch().count();
}

void info(const char *target, const std::string &msg) { printf("INFO  [%s] %s\n", target, msg.c_str()); }

std::vector<float> resolve_betas(const BetaOpt &b, const Opt<float> &lo_beta, const Opt<uint32_t> &lo_num_beta,
                                 bool icm, uint32_t *lo_idx) {
  std::vector<float> betas(b.geometric ? std::max<uint32_t>(b.num_beta, 1) : std::max<size_t>(b.arr.size(), 1));
  uint32_t nb = 0, lo = 0;
  check(tamc_pt_resolve(b.geometric, b.beta_min, b.beta_max, b.num_beta, b.arr.data(), (uint32_t)b.arr.size(),
                        lo_beta.has ? lo_beta.v : NAN, lo_num_beta.has ? (int64_t)lo_num_beta.v : -1, icm,
                        betas.data(), &nb, &lo));
  betas.resize(nb);
  if (lo_idx) *lo_idx = lo;
  return betas;
}

struct Args {
  std::string method, instance, output, sample_output;
  bool has_sample_output = false, qubo = false, dry_run = false;
  std::vector<std::string> suscepts;
};

void usage(FILE *f) {
  fprintf(f,
          "tamc %d\n\nUSAGE:\n    tamc [FLAGS] [OPTIONS] <method-file> <instance-file> <output-file>\n\n"
          "FLAGS:\n    -h, --help       Prints help information\n        --qubo       \n"
          "    -V, --version    Prints version information\n\nOPTIONS:\n"
          "        --sample-output <sample-output>    \n        --suscepts <suscepts>...           \n\n"
          "ARGS:\n    <method-file>      \n    <instance-file>    \n    <output-file>      \n",
          tamc_version());
}

Args parse_args(int argc, char **argv) {
  Args a;
  std::vector<std::string> pos;
  for (int k = 1; k < argc; ++k) {
    std::string s = argv[k];
    if (s == "-h" || s == "--help") {
      usage(stdout);
      exit(0);
    } else if (s == "-V" || s == "--version") {
      printf("tamc %d\n", tamc_version());
      exit(0);
    } else if (s == "--qubo") {
      a.qubo = true;
    } else if (s == "--dry-run") {
      a.dry_run = true;
    } else if (s == "--sample-output" || s.rfind("--sample-output=", 0) == 0) {
      if (s.size() > 15 && s[15] == '=') {
        a.sample_output = s.substr(16);
      } else {
        if (k + 1 >= argc) throw Error("The argument '--sample-output <sample-output>' requires a value");
        a.sample_output = argv[++k];
      }
      a.has_sample_output = true;
    } else if (s == "--suscepts" || s.rfind("--suscepts=", 0) == 0) {
      if (s.size() > 10 && s[10] == '=') {
        a.suscepts.push_back(s.substr(11));
      } else {
        // takes values up to the next option (structopt Vec<String>)
        int got = 0;
        while (k + 1 < argc && argv[k + 1][0] != '-') {
          a.suscepts.push_back(argv[++k]);
          ++got;
        }
        if (!got) throw Error("The argument '--suscepts <suscepts>...' requires a value");
      }
    } else if (s == "--") {
      for (++k; k < argc; ++k) pos.push_back(argv[k]);
    } else if (s.size() > 1 && s[0] == '-') {
      throw Error("Found argument '" + s + "' which wasn't expected");
    } else {
      pos.push_back(s);
    }
  }
  // --suscepts may have swallowed positionals given after it: the last three values are the
  // positional arguments when too few were seen (structopt resolves the same way)
  while (pos.size() < 3 && !a.suscepts.empty()) {
    pos.insert(pos.begin(), a.suscepts.back());
    a.suscepts.pop_back();
  }
  if (pos.size() != 3) throw Error("expected <method-file> <instance-file> <output-file>");
  a.method = pos[0];
  a.instance = pos[1];
  a.output = pos[2];
  return a;
}

std::string read_file(const std::string &path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw Error("IO Error (" + path + ")");
  std::stringstream ss;
  ss << f.rdbuf();
  return ss.str();
}

std::string json_floats(const std::vector<float> &v) {
  std::string s = "[";
  for (size_t k = 0; k < v.size(); ++k) s += (k ? ", " : "") + yaml_float(v[k], true);
  return s + "]";
}

int run_pt(const Args &a, const PtMethod &p) {
  const std::string sample_output = a.has_sample_output ? a.sample_output : "samples.bin";
  tamc_instance *inst = nullptr;
  check(tamc_instance_from_file(a.instance.c_str(), a.qubo, &inst));
  std::unique_ptr<tamc_instance, void (*)(tamc_instance *)> guard(inst, tamc_instance_destroy);
  for (auto &f : a.suscepts) check(tamc_instance_add_suscept_file(inst, f.c_str()));
  uint32_t lo = 0;
  const std::vector<float> betas = resolve_betas(p.beta, p.lo_beta, p.lo_num_beta, p.icm, &lo);
  tamc_instance_info ii{};
  check(tamc_instance_get_info(inst, &ii));
  if (a.dry_run) {
    printf("{\"method\": \"PT\", \"n\": %u, \"nnz\": %u, \"nchi\": %u, \"qubo\": %d, \"num_sweeps\": %u, "
           "\"betas\": %s, \"lo_beta_idx\": %u, \"icm\": %s, \"num_replica_chains\": %u, \"threads\": %u, "
           "\"sample\": %s, \"sample_states\": %s, \"sample_limiting\": %s, \"output\": \"%s\", "
           "\"sample_output\": \"%s\", \"sample_format\": \"%s\"}\n",
           ii.n, ii.nnz, ii.nchi, ii.qubo, p.num_sweeps, json_floats(betas).c_str(), lo, p.icm ? "true" : "false",
           p.num_replica_chains, p.threads, p.sample.has ? std::to_string(p.sample.v).c_str() : "null",
           p.sample_states.has ? std::to_string(p.sample_states.v).c_str() : "null",
           p.sample_limiting.has ? std::to_string(p.sample_limiting.v).c_str() : "null", a.output.c_str(),
           sample_output.c_str(), ends_with(sample_output, "pkl") ? "pickle" : "bincode");
    return 0;
  }
  check(tamc_init());
  printf(" ** Parallel Tempering - ICM **\n");
  info("tamc::pt", "Number of sweeps: " + std::to_string(p.num_sweeps));
  info("tamc::pt", p.icm ? "Using ICM" : "ICM Disabled");
  tamc_pt_params P{};
  P.num_sweeps = p.num_sweeps;
  P.warmup_fraction = p.warmup_fraction;
  P.num_betas = (uint32_t)betas.size();
  P.betas = betas.data();
  P.lo_beta_idx = lo;
  P.icm = p.icm;
  P.num_replica_chains = p.num_replica_chains;
  P.threads = p.threads;
  P.sample_enabled = p.sample.has;
  P.sample = p.sample.v;
  P.sample_states_enabled = p.sample_states.has;
  P.sample_states = p.sample_states.v;
  P.sample_limiting = p.sample_limiting.has ? p.sample_limiting.v : 0;
  P.seed = run_seed();
  tamc_pt_result *r = nullptr;
  check(tamc_pt_run(inst, &P, nullptr, &r));
  std::unique_ptr<tamc_pt_result, void (*)(tamc_pt_result *)> rg(r, tamc_pt_result_free);
  printf("PT-ICM Done.\n** Ground state energy **\n");
  if (r->n_gs == 0) throw Error("no sweep after warm-up: the ground-state trace is empty");
  printf("  e = %s\n", rust_display(r->gs_energies[r->n_gs - 1], true).c_str());
  const uint32_t nsw = r->gs_steps[r->n_gs - 1];
  printf("  nsw = %u\n", nsw);
  printf("  %% remaining = %4.3f%%\n", (double)(p.num_sweeps - nsw) / (double)p.num_sweeps * 100.0);
  write_text(a.output, pt_min_results_yaml(p, *r), "yaml output file");
  if (ends_with(sample_output, "pkl")) check(tamc_write_thermal_samples_pickle(r, betas.data(), 0, sample_output.c_str()));
  else check(tamc_write_thermal_samples(r, betas.data(), 0, sample_output.c_str()));
  return 0;
}

int run_sa(const Args &a, const SaMethod &m) {
  tamc_instance *inst = nullptr;
  check(tamc_instance_from_file(a.instance.c_str(), a.qubo, &inst));
  std::unique_ptr<tamc_instance, void (*)(tamc_instance *)> guard(inst, tamc_instance_destroy);
  Opt<float> no_lo;
  Opt<uint32_t> no_lon;
  const std::vector<float> betas = resolve_betas(m.beta, no_lo, no_lon, false, nullptr);
  const std::string sample_output = a.has_sample_output ? a.sample_output : "samples.bin";
  if (a.dry_run) {
    tamc_instance_info ii{};
    check(tamc_instance_get_info(inst, &ii));
    printf("{\"method\": \"SA\", \"n\": %u, \"nnz\": %u, \"qubo\": %d, \"betas\": %s, \"num_replicas\": %u, "
           "\"threads\": %u, \"output\": \"%s\", \"sample_output\": \"%s\"}\n",
           ii.n, ii.nnz, ii.qubo, json_floats(betas).c_str(), m.num_replicas, m.threads, a.output.c_str(),
           sample_output.c_str());
    return 0;
  }
  check(tamc_init());
  info("tamc::sa", " ** Simulated Annealing **");
  if (m.num_replicas > 1) info("tamc::sa", "Running multi-threaded SA");
  tamc_sa_params P{};
  P.num_betas = (uint32_t)betas.size();
  P.betas = betas.data();
  P.num_replicas = m.num_replicas;
  P.threads = m.threads;
  P.seed = run_seed();
  tamc_sa_result *r = nullptr;
  check(tamc_sa_run(inst, &P, &r));
  std::unique_ptr<tamc_sa_result, void (*)(tamc_sa_result *)> rg(r, tamc_sa_result_free);
  uint32_t ngs = 0;
  for (uint32_t k = 0; k < r->R; ++k) ngs += r->energies[k] <= r->min_energy;
  const double pgs = (double)ngs / (double)r->R;
  const double tts = (r->timing_us / 1e6 / (double)m.num_replicas) * (log10(0.01) / log10(1.0 - pgs));
  char buf[512];
  snprintf(buf, sizeof buf, "\n** Ground state energy **\n  e = %s\n** GS Probability  **\n  p_gs = %.4e\n"
           "** Time-to-solution **\n  tts = %s s\n",
           rust_display(r->min_energy, true).c_str(), pgs, rust_display(tts, false).c_str());
  info("tamc::sa", buf);
  if (ends_with(a.output, "pkl")) write_text(a.output, sa_min_results_pickle(m, *r, r->timing_us), "output file");
  else if (ends_with(a.output, "yml")) write_text(a.output, sa_min_results_yaml(m, *r, r->timing_us), "output file");
  if (ends_with(sample_output, "pkl")) check(tamc_write_anneal_state_pickle(r, sample_output.c_str()));
  else check(tamc_write_anneal_state(r, sample_output.c_str()));
  return 0;
}

tamc_pt_params pt_params_of(const PtMethod &p, const std::vector<float> &betas, uint32_t lo) {
  tamc_pt_params P{};
  P.num_sweeps = p.num_sweeps;
  P.warmup_fraction = p.warmup_fraction;
  P.num_betas = (uint32_t)betas.size();
  P.betas = betas.data();
  P.lo_beta_idx = lo;
  P.icm = p.icm;
  P.num_replica_chains = p.num_replica_chains;
  P.threads = p.threads;
  P.sample_enabled = p.sample.has;
  P.sample = p.sample.v;
  P.sample_states_enabled = p.sample_states.has;
  P.sample_states = p.sample_states.v;
  P.sample_limiting = p.sample_limiting.has ? p.sample_limiting.v : 0;
  return P;
}

// run_gla (src/gla.rs:122-196): the partitions' sub-instances are solved in one batched launch.
int run_gla(const Args &a, const GlaMethod &g) {
  tamc_instance *inst = nullptr;
  check(tamc_instance_from_file(a.instance.c_str(), a.qubo, &inst));
  std::unique_ptr<tamc_instance, void (*)(tamc_instance *)> guard(inst, tamc_instance_destroy);
  tamc_partition *parts = nullptr;
  check(tamc_partition_from_file(inst, g.partition_file.c_str(), &parts));
  std::unique_ptr<tamc_partition, void (*)(tamc_partition *)> pguard(parts, tamc_partition_free);
  uint32_t lo = 0;
  const std::vector<float> betas = resolve_betas(g.pt.beta, g.pt.lo_beta, g.pt.lo_num_beta, g.pt.icm, &lo);
  if (a.dry_run) {
    printf("{\"method\": \"GLA\", \"n\": %u, \"num_partitions\": %u, \"num_threads\": %u, \"num_sweeps\": %u, "
           "\"betas\": %s, \"lo_beta_idx\": %u, \"icm\": %s, \"output\": \"%s\", \"output_format\": \"%s\"}\n",
           parts->num_partitions ? parts->offsets[parts->num_partitions] : 0u, parts->num_partitions, g.num_threads,
           g.pt.num_sweeps, json_floats(betas).c_str(), lo, g.pt.icm ? "true" : "false", a.output.c_str(),
           ends_with(a.output, "pkl") ? "pickle" : "bincode");
    return 0;
  }
  check(tamc_init());
  info("tamc::gla", " ** Running GLA with PT-ICM sub-solvers ...");
  tamc_pt_params P = pt_params_of(g.pt, betas, lo);
  P.seed = run_seed();
  tamc_gla_result *r = nullptr;
  check(tamc_gla_run(inst, parts, &P, &r));
  std::unique_ptr<tamc_gla_result, void (*)(tamc_gla_result *)> rg(r, tamc_gla_result_free);
  info("tamc::gla", "Bulk energy: " + rust_display(r->bulk_energy, true));
  info("tamc::gla", "Boundary energy: " + rust_display(r->boundary_energy, true));
  info("tamc::gla", "Total energy: " + rust_display(r->total_energy, true));
  check(tamc_write_gla_results(r, a.output.c_str()));
  return 0;
}

int run(int argc, char **argv) {
  const Args a = parse_args(argc, argv);
  const std::string text = read_file(a.method);
  Node doc;
  try {
    doc = Yaml(text).parse();
  } catch (const Error &e) {
    throw Error("Failed to parse method file (" + a.method + "): " + e.what());
  }
  if (doc.kind != Node::Map || doc.map.size() != 1)
    throw Error("Failed to parse method file (" + a.method + "): expected one of `PT`, `SA`, `GLA`");
  const std::string &tag = doc.map[0].first;
  try {
    if (tag == "PT") return run_pt(a, parse_pt(doc.map[0].second));
    if (tag == "SA") return run_sa(a, parse_sa(doc.map[0].second));
    if (tag == "GLA") return run_gla(a, parse_gla(doc.map[0].second));
  } catch (const Error &e) {
    const std::string w = e.what();
    if (w.rfind("PT:", 0) == 0 || w.rfind("SA:", 0) == 0 || w.rfind("GLA:", 0) == 0 || w.rfind("invalid", 0) == 0 ||
        w.rfind("`", 0) == 0)
      throw Error("Failed to parse method file (" + a.method + "): " + w);
    throw;
  }
  throw Error("Failed to parse method file (" + a.method + "): unknown variant `" + tag +
              "`, expected one of `PT`, `SA`, `GLA`");
}

// ------------------------------------------------------------------ tamc-pt-opt ----
// src/bin/tamc-pt-opt.rs: optimise the PT temperature set over a set of instances.  Every iteration
// runs PT on all instances in one batched launch (iteration i, instance k: seed + i * count + k),
// starting from the previous iteration's final states, and moves the set towards equal diffusion
// steps (tamc_pt_opt_update).
struct OptArgs {
  uint32_t max_iters = 1000;
  float step_size = 0.2f, momentum = 0.85f, tolerance = 0.02f;
  std::string opt_params = "opt_params.yml", tau_hist = "tau_hist.csv", params;
  std::vector<std::string> instances;
};

OptArgs parse_opt_args(int argc, char **argv) {
  OptArgs o;
  std::vector<std::string> pos;
  auto value = [&](int &k, const std::string &s, const char *name) -> std::string {
    const size_t eq = s.find('=');
    if (eq != std::string::npos) return s.substr(eq + 1);
    if (k + 1 >= argc) throw Error(std::string("The argument '") + name + "' requires a value");
    return argv[++k];
  };
  for (int k = 1; k < argc; ++k) {
    const std::string s = argv[k];
    const std::string key = s.substr(0, s.find('='));
    if (s == "-h" || s == "--help") {
      printf("pt-opt\nOptimize the PT temperature set over a set of instances.\n\nUSAGE:\n    tamc-pt-opt [OPTIONS] "
             "<params> [instances]...\n\nOPTIONS:\n        --max-iters <max-iters>       [default: 1000]\n"
             "        --momentum <momentum>         [default: 0.85]\n        --opt-params <opt-params>     "
             "Destination for optimized parameters [default: opt_params.yml]\n        --step-size <step-size>       "
             "[default: 0.2]\n        --tau-hist <tau-hist>         [default: tau_hist.csv]\n        --tolerance "
             "<tolerance>       Covergence tolerance criterion (MAE of log-betas) [default: 0.02]\n");
      exit(0);
    } else if (key == "--max-iters") {
      o.max_iters = (uint32_t)std::stoul(value(k, s, "--max-iters <max-iters>"));
    } else if (key == "--step-size") {
      o.step_size = std::stof(value(k, s, "--step-size <step-size>"));
    } else if (key == "--momentum") {
      o.momentum = std::stof(value(k, s, "--momentum <momentum>"));
    } else if (key == "--tolerance") {
      o.tolerance = std::stof(value(k, s, "--tolerance <tolerance>"));
    } else if (key == "--opt-params") {
      o.opt_params = value(k, s, "--opt-params <opt-params>");
    } else if (key == "--tau-hist") {
      o.tau_hist = value(k, s, "--tau-hist <tau-hist>");
    } else if (s.size() > 1 && s[0] == '-') {
      throw Error("Found argument '" + s + "' which wasn't expected");
    } else {
      pos.push_back(s);
    }
  }
  if (pos.empty()) throw Error("The following required arguments were not provided: <params>");
  o.params = pos[0];
  o.instances.assign(pos.begin() + 1, pos.end());
  return o;
}

std::string fmt(const char *f, double v) {
  char b[64];
  snprintf(b, sizeof b, f, v);
  return b;
}

int run_pt_opt(int argc, char **argv) {
  const OptArgs o = parse_opt_args(argc, argv);
  std::vector<std::unique_ptr<tamc_instance, void (*)(tamc_instance *)>> insts;
  std::vector<const tamc_instance *> ptrs;
  for (auto &f : o.instances) {
    tamc_instance *inst = nullptr;
    check(tamc_instance_from_file(f.c_str(), 0, &inst));
    insts.emplace_back(inst, tamc_instance_destroy);
    ptrs.push_back(inst);
  }
  const Node doc = Yaml(read_file(o.params)).parse();
  if (doc.kind != Node::Map || doc.map.size() != 1 || doc.map[0].first != "PT") {
    fprintf(stderr, "WARN  [tamc_pt_opt] Must input a PT method for tamc-pt-opt\n");
    return 0;
  }
  PtMethod p = parse_pt(doc.map[0].second);
  if (ptrs.empty()) throw Error("no instances given");
  check(tamc_init());
  const uint32_t count = (uint32_t)ptrs.size(), nc = p.num_replica_chains;
  uint32_t lo = 0;
  std::vector<float> betas = resolve_betas(p.beta, p.lo_beta, p.lo_num_beta, p.icm, &lo);
  const uint32_t nt = (uint32_t)betas.size();
  std::vector<float> momentum = betas;
  std::vector<std::vector<float>> tau_hist;
  std::vector<std::vector<int8_t>> init(count);
  const uint64_t seed = run_seed();
  for (uint32_t it = 0; it < o.max_iters; ++it) {
    printf("* Iteration %u.\n* Step Size: %s\n", it, rust_display(o.step_size, true).c_str());
    betas = resolve_betas(p.beta, p.lo_beta, p.lo_num_beta, p.icm, &lo);
    std::vector<tamc_pt_params> P(count);
    std::vector<const int8_t *> init_ptrs(count, nullptr);
    for (uint32_t k = 0; k < count; ++k) {
      P[k] = pt_params_of(p, betas, lo);
      P[k].seed = seed + (uint64_t)it * count + k;
      if (!init[k].empty()) init_ptrs[k] = init[k].data();
    }
    std::vector<tamc_pt_result *> res(count, nullptr);
    tamc_batch_options bo{};
    tamc_batch_stats bs{};
    check(tamc_pt_run_batch_init(count, ptrs.data(), P.data(), init_ptrs.data(), &bo, res.data(), &bs));
    struct Free {
      std::vector<tamc_pt_result *> &r;
      ~Free() {
        for (auto *x : r) tamc_pt_result_free(x);
      }
    } fr{res};
    std::vector<uint32_t> hist((size_t)count * nt * 2, 0u), rts(count, 0u);
    for (uint32_t k = 0; k < count; ++k) {
      const tamc_pt_result *r = res[k];
      for (uint32_t c = 0; c < r->C; ++c) {
        rts[k] += r->round_trips[c];
        for (uint32_t b = 0; b < nt * 2; ++b) hist[(size_t)k * nt * 2 + b] += r->diffusion_hist[((size_t)c * nt) * 2 + b];
      }
    }
    std::vector<float> next(nt), taus(count);
    float err = 0.f;
    int updated = 0;
    check(tamc_pt_opt_update(nt, betas.data(), momentum.data(), count, hist.data(), rts.data(), p.num_sweeps, nc,
                             o.step_size, o.momentum, next.data(), &err, taus.data(), &updated));
    std::string tl = "Round-trip times (sweeps):\n[";
    for (uint32_t k = 0; k < count; ++k) tl += (k ? ", " : "") + rust_display(taus[k], true);
    printf("%s]\n", (tl).c_str());
    tau_hist.push_back(taus);
    if (!updated) {
      fprintf(stderr, "WARN  [tamc_pt_opt]  ** Insufficient round trips for eta CDF\n");
      continue;
    }
    printf("%9s%9s\n--------------------------------------------------\n", "beta0", "beta");
    for (uint32_t b = 0; b < nt; ++b) printf(" %7.4f  %7.4f \n", betas[b], next[b]);
    printf("Mean abs log rel_err: %s\n", rust_display(err, true).c_str());
    p.beta.geometric = false;
    p.beta.arr = next;
    for (uint32_t k = 0; k < count; ++k) {
      const tamc_pt_result *r = res[k];
      init[k].assign(r->final_states, r->final_states + (size_t)r->C * r->B * r->n);
    }
    if (err < o.tolerance) {
      printf(" ** Relative Error converged\n");
      break;
    }
  }
  betas = resolve_betas(p.beta, p.lo_beta, p.lo_num_beta, p.icm, &lo);
  std::string fl = "Final temperature array:\n[";
  for (uint32_t b = 0; b < nt; ++b) fl += (b ? ", " : "") + fmt("%6.5f", betas[b]);
  printf("%s]\n", fl.c_str());
  p.num_sweeps *= 2;
  YamlOut y;
  y.line(0, "PT:");
  pt_params_yaml(y, 2, p);
  printf("%s\n", y.s.c_str());
  printf("\n\n ** Writing to %s **\n", o.opt_params.c_str());
  write_text(o.opt_params, y.s, "optimized parameters");
  std::string csv;
  for (auto &row : tau_hist) {
    for (float x : row) csv += rust_display(x, true) + ",";
    csv += "\n";
  }
  write_text(o.tau_hist, csv, "tau history");
  return 0;
}

// ------------------------------------------------------------------ percolate ----
// src/bin/percolate.rs: site percolation samples of an instance file's graph.
int run_percolate(int argc, char **argv) {
  uint32_t samples = 100;
  double prob = -1.0;
  bool has_prob = false;
  std::string output, instance;
  bool has_output = false, has_instance = false;
  for (int k = 1; k < argc; ++k) {
    const std::string s = argv[k];
    auto value = [&](const char *name) -> std::string {
      const size_t eq = s.find('=');
      if (s.rfind("--", 0) == 0 && eq != std::string::npos) return s.substr(eq + 1);
      if (k + 1 >= argc) throw Error(std::string("The argument '") + name + "' requires a value");
      return argv[++k];
    };
    const std::string key = s.substr(0, s.find('='));
    if (s == "-h" || s == "--help") {
      printf("percolate\nPerform a site percolation calculation on an arbitrary graph.\n\nUSAGE:\n    percolate "
             "[OPTIONS] <instance> --prob <prob>\n\nOPTIONS:\n    -n, --num-samples <num-samples>     [default: 100]\n"
             "    -o, --output <output>               \n    -p, --prob <prob>                   \n");
      exit(0);
    } else if (key == "-n" || key == "--num-samples") {
      samples = (uint32_t)std::stoul(value("--num-samples <num-samples>"));
    } else if (key == "-p" || key == "--prob") {
      prob = std::stod(value("--prob <prob>"));
      has_prob = true;
    } else if (key == "-o" || key == "--output") {
      output = value("--output <output>");
      has_output = true;
    } else if (s.size() > 1 && s[0] == '-') {
      throw Error("Found argument '" + s + "' which wasn't expected");
    } else if (!has_instance) {
      instance = s;
      has_instance = true;
    } else {
      throw Error("Found argument '" + s + "' which wasn't expected");
    }
  }
  if (!has_prob || !has_instance)
    throw Error("The following required arguments were not provided: --prob <prob> <instance>");
  tamc_instance *inst = nullptr;
  check(tamc_instance_from_file(instance.c_str(), 0, &inst));
  std::unique_ptr<tamc_instance, void (*)(tamc_instance *)> guard(inst, tamc_instance_destroy);
  check(tamc_init());
  std::vector<uint32_t> cl(std::max<uint32_t>(samples, 1));
  std::vector<uint8_t> pc(std::max<uint32_t>(samples, 1));
  check(tamc_percolate(inst, prob, samples, run_seed(), cl.data(), pc.data()));
  uint32_t num = 0;
  for (uint32_t k = 0; k < samples; ++k) num += pc[k];
  printf("# * Percolation results *\n");
  printf("samps,  num_perc,  frac_perc\n");
  printf("%u, %u, %s\n", samples, num, rust_display((double)num / (double)samples, false).c_str());
  if (has_output) {
    // the reference writes the header and then every sample's largest cluster with no separator
    std::string text = "perc, max_cluster";
    for (uint32_t k = 0; k < samples; ++k) text += std::to_string(cl[k]);
    write_text(output, text, "output file");
  }
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  {
    const char *prog = strrchr(argv[0], '/');
    prog = prog ? prog + 1 : argv[0];
    if (strcmp(prog, "percolate") == 0) {
      try {
        return run_percolate(argc, argv);
      } catch (const std::exception &e) {
        fflush(stdout);
        fprintf(stderr, "percolate terminated with an error:\n%s\n", e.what());
        return 1;
      }
    }
    if (strcmp(prog, "tamc-pt-opt") == 0) {
      try {
        return run_pt_opt(argc, argv);
      } catch (const std::exception &e) {
        fflush(stdout);
        fprintf(stderr, "tamc-pt-opt terminated with an error:\n%s\n", e.what());
        return 1;
      }
    }
  }
  try {
    const int rc = run(argc, argv);
    bool dry = false;
    for (int k = 1; k < argc; ++k) dry |= strcmp(argv[k], "--dry-run") == 0;
    if (!dry) {
      printf("tamc finished successfully\n");
      fprintf(stderr, " * * * * * * \n");
    }
    return rc;
  } catch (const std::exception &e) {
    fflush(stdout);
    fprintf(stderr, "tamc terminated with an error:\n%s\n * * * * * * \n", e.what());
    return 1;
  }
}
