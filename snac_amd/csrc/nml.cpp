// Fortran namelist reader for SNaC's dns.nml / blocks.nml (param.f90:88-106).
// Supports &group ... / blocks, `name = v`, `name(i:j,k,...) = v1, v2, ...`
// (filled with the first index fastest, as Fortran does), quoted strings,
// logicals (T/F/.true./.false.), numbers with e/d exponents, `n*value`
// repeats and `!` comments.
#include <algorithm>
#include <cctype>
#include <fstream>
#include <sstream>
#include <stdexcept>

#include "case.hpp"

namespace snac {
namespace {

struct Tok {
  enum K { Word, Str, Sym } k;
  std::string s;
};

std::vector<Tok> tokenize(const std::string &t) {
  std::vector<Tok> out;
  size_t i = 0, n = t.size();
  while (i < n) {
    char c = t[i];
    if (c == ' ' || c == '\t' || c == '\r' || c == '\n' || c == ',') { i++; continue; }
    if (c == '!') { while (i < n && t[i] != '\n') i++; continue; }
    if (c == '\'' || c == '"') {
      size_t j = t.find(c, i + 1);
      if (j == std::string::npos) throw std::runtime_error("unterminated string in namelist");
      out.push_back({Tok::Str, t.substr(i + 1, j - i - 1)});
      i = j + 1;
      continue;
    }
    if (c == '&' || c == '/' || c == '=' || c == '(' || c == ')') {
      out.push_back({Tok::Sym, std::string(1, c)});
      i++;
      continue;
    }
    size_t j = i;
    while (j < n && std::string(" \t\r\n,=()!&/'\"").find(t[j]) == std::string::npos) j++;
    out.push_back({Tok::Word, t.substr(i, j - i)});
    i = j;
  }
  return out;
}

std::string lower(std::string s) {
  for (auto &c : s) c = (char)std::tolower((unsigned char)c);
  return s;
}

bool is_number(const std::string &s) {
  if (s.empty()) return false;
  bool digit = false;
  for (char c : s) {
    if (std::isdigit((unsigned char)c)) digit = true;
    else if (std::string(".eEdD+-").find(c) == std::string::npos) return false;
  }
  return digit;
}

NmlValue to_value(const Tok &tk) {
  NmlValue v;
  if (tk.k == Tok::Str) { v.kind = NmlValue::Str; v.str = tk.s; return v; }
  std::string l = lower(tk.s);
  if (l == "t" || l == ".t." || l == ".true." || l == "true") { v.kind = NmlValue::Bool; v.b = true; return v; }
  if (l == "f" || l == ".f." || l == ".false." || l == "false") { v.kind = NmlValue::Bool; v.b = false; return v; }
  if (is_number(l)) {
    for (auto &c : l) if (c == 'd') c = 'e';
    v.kind = NmlValue::Num;
    v.num = std::stod(l);
    return v;
  }
  v.kind = NmlValue::Str;
  v.str = tk.s;
  return v;
}

// an open bound (`alphai(:)`) starts at 1 and extends over the nvals values
std::vector<std::vector<int>> expand_index(const std::vector<std::string> &parts, int nvals) {
  std::vector<std::pair<int, int>> rng;
  for (const auto &p : parts) {
    size_t c = p.find(':');
    if (c == std::string::npos) { int v = std::stoi(p); rng.push_back({v, v}); continue; }
    const std::string a = p.substr(0, c), b = p.substr(c + 1);
    const int lo = a.empty() ? 1 : std::stoi(a);
    const int hi = b.empty() ? lo + std::max(nvals, 1) - 1 : std::stoi(b);
    rng.push_back({lo, hi});
  }
  std::vector<std::vector<int>> out;
  std::vector<int> cur(rng.size());
  for (size_t d = 0; d < rng.size(); d++) cur[d] = rng[d].first;
  if (rng.empty()) return out;
  for (;;) {
    out.push_back(cur);
    size_t d = 0;  // first index fastest
    while (d < rng.size()) {
      if (++cur[d] <= rng[d].second) break;
      cur[d] = rng[d].first;
      d++;
    }
    if (d == rng.size()) break;
  }
  return out;
}

}  // namespace

Nml read_namelist_file(const std::string &path) {
  std::ifstream f(path);
  if (!f) throw std::runtime_error("cannot open " + path);
  std::stringstream ss;
  ss << f.rdbuf();
  auto toks = tokenize(ss.str());
  Nml out;
  NmlGroup *cur = nullptr;
  size_t i = 0;
  while (i < toks.size()) {
    const Tok &t = toks[i];
    if (t.k == Tok::Sym && t.s == "&") {
      if (i + 1 >= toks.size()) break;
      cur = &out[lower(toks[i + 1].s)];
      i += 2;
      continue;
    }
    if (t.k == Tok::Sym && t.s == "/") { cur = nullptr; i++; continue; }
    if (!cur || t.k != Tok::Word) { i++; continue; }
    std::string name = lower(t.s);
    i++;
    std::vector<std::string> idx;
    bool has_idx = false;
    if (i < toks.size() && toks[i].k == Tok::Sym && toks[i].s == "(") {
      has_idx = true;
      i++;
      while (i < toks.size() && !(toks[i].k == Tok::Sym && toks[i].s == ")")) idx.push_back(toks[i++].s);
      i++;
    }
    if (i >= toks.size() || !(toks[i].k == Tok::Sym && toks[i].s == "="))
      throw std::runtime_error("namelist " + path + ": expected '=' after " + name);
    i++;
    std::vector<NmlValue> vals;
    while (i < toks.size()) {
      const Tok &v = toks[i];
      if (v.k == Tok::Sym) break;
      if (v.k == Tok::Word && i + 1 < toks.size() && toks[i + 1].k == Tok::Sym &&
          (toks[i + 1].s == "=" || toks[i + 1].s == "(") && !is_number(v.s))
        break;
      size_t star = v.s.find('*');
      if (v.k == Tok::Word && star != std::string::npos && star > 0 &&
          std::isdigit((unsigned char)v.s[0])) {
        int cnt = std::stoi(v.s.substr(0, star));
        NmlValue val = to_value({Tok::Word, v.s.substr(star + 1)});
        for (int q = 0; q < cnt; q++) vals.push_back(val);
      } else {
        vals.push_back(to_value(v));
      }
      i++;
    }
    auto &entry = (*cur)[name];
    if (!has_idx) {
      if (vals.size() == 1) entry[{}] = vals[0];
      else for (size_t q = 0; q < vals.size(); q++) entry[{(int)q + 1}] = vals[q];
    } else {
      auto combos = expand_index(idx, (int)vals.size());
      for (size_t q = 0; q < vals.size() && q < combos.size(); q++) entry[combos[q]] = vals[q];
    }
  }
  return out;
}

}  // namespace snac
