// f2m_tsplib.cpp -- TSPLIB coordinate reader (host).
//
// Replaces get_size + FXY::get_KXY (FXY.hpp:186-309): lines before
// NODE_COORD_SECTION are ignored, reading stops at a line starting with EOF,
// the data is read as a stream of `label x y` triples; x and y are rounded to
// 1e-3 (std::round(v*1000)/1000, FXY.hpp:252-256) and a point equal to an
// earlier one is dropped and recorded against the earlier label
// (FXY.hpp:262-270).  The reference copies the data into `<file>_tmp` first;
// this reader parses in memory and writes nothing next to the input.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <unordered_map>
#include <vector>

#include "f2m.h"

struct f2m_points {
  std::vector<int64_t> label;
  std::vector<double> x, y;
  std::map<int64_t, std::vector<int64_t>> dups;  // kept label -> dropped labels
  int64_t n_dup = 0;
};

namespace {
// Hash of a rounded point.  -0.0 == +0.0 under the map's equality (and
// std::hash<double> gives both the same value, as the reference's
// unordered_map<double, ...> relies on, FXY.hpp:259-267), so a zero is
// hashed as +0.0: a point rounding to -0.0 then meets an earlier +0.0.
struct PairHash {
  size_t operator()(const std::pair<double, double>& p) const {
    const double x = p.first == 0.0 ? 0.0 : p.first;
    const double y = p.second == 0.0 ? 0.0 : p.second;
    uint64_t a, b;
    std::memcpy(&a, &x, 8);
    std::memcpy(&b, &y, 8);
    return std::hash<uint64_t>()(a * 0x9E3779B97F4A7C15ull ^ (b + 0x7F4A7C15ull));
  }
};
}  // namespace

extern "C" {

f2m_status f2m_tsplib_read(const char* path, f2m_points** out) {
  if (!path || !out) return F2M_ERR_ARG;
  *out = nullptr;
  FILE* f = std::fopen(path, "rb");
  if (!f) return F2M_ERR_IO;
  std::string text;
  {
    char buf[1 << 16];
    size_t r;
    while ((r = std::fread(buf, 1, sizeof buf, f)) > 0) text.append(buf, r);
    std::fclose(f);
  }
  // locate the data block: after the NODE_COORD_SECTION line, up to EOF line
  size_t pos = 0, data_begin = std::string::npos, data_end = text.size();
  while (pos < text.size()) {
    size_t eol = text.find('\n', pos);
    if (eol == std::string::npos) eol = text.size();
    const char* line = text.data() + pos;
    size_t len = eol - pos;
    if (len >= 3 && std::strncmp(line, "EOF", 3) == 0) {
      if (data_begin != std::string::npos) data_end = pos;
      else data_end = pos;
      break;
    }
    if (data_begin == std::string::npos && len >= 18 &&
        std::strncmp(line, "NODE_COORD_SECTION", 18) == 0)
      data_begin = eol < text.size() ? eol + 1 : eol;
    pos = eol + 1;
  }
  if (data_begin == std::string::npos || data_begin >= data_end) return F2M_ERR_FORMAT;
  std::string data = text.substr(data_begin, data_end - data_begin);

  f2m_points* P = new f2m_points;
  std::unordered_map<std::pair<double, double>, int64_t, PairHash> seen;
  const char* s = data.c_str();
  const char* e = s + data.size();
  while (s < e) {
    char* q;
    long long k = std::strtoll(s, &q, 10);
    if (q == s) break;
    s = q;
    double x = std::strtod(s, &q);
    if (q == s) break;
    s = q;
    double y = std::strtod(s, &q);
    if (q == s) break;
    s = q;
    x = std::round(x * 1000) / 1000;
    y = std::round(y * 1000) / 1000;
    if (k < 0) continue;  // FXY.hpp:258 keeps only labels > -1
    auto it = seen.find({x, y});
    if (it != seen.end()) {
      P->dups[it->second].push_back(k);
      P->n_dup++;
      continue;
    }
    seen.emplace(std::make_pair(x, y), (int64_t)k);
    P->label.push_back(k);
    P->x.push_back(x);
    P->y.push_back(y);
  }
  if (P->label.empty()) {
    delete P;
    return F2M_ERR_FORMAT;
  }
  *out = P;
  return F2M_OK;
}

int64_t f2m_points_count(const f2m_points* p) { return p ? (int64_t)p->label.size() : 0; }
int64_t f2m_points_duplicates(const f2m_points* p) { return p ? p->n_dup : 0; }

void f2m_points_copy(const f2m_points* p, int64_t* labels, double* x, double* y) {
  if (!p) return;
  size_t n = p->label.size();
  if (labels) std::memcpy(labels, p->label.data(), n * sizeof(int64_t));
  if (x) std::memcpy(x, p->x.data(), n * sizeof(double));
  if (y) std::memcpy(y, p->y.data(), n * sizeof(double));
}

// FXY::save_to_file_duplicates (FXY.hpp:168-183), kept labels in ascending order.
f2m_status f2m_points_write_duplicates(const f2m_points* p, const char* path) {
  if (!p || !path) return F2M_ERR_ARG;
  FILE* f = std::fopen(path, "w");
  if (!f) return F2M_ERR_IO;
  for (auto& kv : p->dups) {
    std::fprintf(f, "node used: %lld duplicates: ", (long long)kv.first);
    for (int64_t d : kv.second) std::fprintf(f, "%lld ", (long long)d);
    std::fprintf(f, "\n");
  }
  std::fclose(f);
  return F2M_OK;
}

void f2m_points_free(f2m_points* p) { delete p; }

}  // extern "C"
