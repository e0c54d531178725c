/*
 * harness_common.hpp -- parameter-file and raw-state I/O shared by the
 * reference harnesses (TEST INFRASTRUCTURE).  The harnesses subclass the
 * reference's own driver classes (the extension mechanism the reference uses
 * itself, cf. 2D/SERIAL/MRX/dmhd2d_class.hpp:8) to reach the protected
 * ideal()/bound()/setdt() and run an exact number of steps on a state read
 * from disk.  Doubles in the parameter file are C99 hex floats (exact).
 */
#ifndef QAS_HARNESS_COMMON_HPP
#define QAS_HARNESS_COMMON_HPP
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <map>
#include <vector>
#include <sstream>
#include <fstream>

struct HParams {
  std::map<std::string, std::string> kv;
  bool load(const char *path) {
    std::ifstream f(path);
    if (!f) return false;
    std::string line;
    while (std::getline(f, line)) {
      std::istringstream ss(line);
      std::string key;
      if (!(ss >> key)) continue;
      std::string rest;
      std::getline(ss, rest);
      size_t p = rest.find_first_not_of(" \t");
      kv[key] = (p == std::string::npos) ? "" : rest.substr(p);
    }
    return true;
  }
  std::string s(const char *k) const { auto it = kv.find(k); return it == kv.end() ? "" : it->second; }
  double d(const char *k) const { return std::strtod(s(k).c_str(), nullptr); }
  int i(const char *k) const { return std::atoi(s(k).c_str()); }
  void ints(const char *k, int *out, int n) const {
    std::istringstream ss(s(k));
    for (int q = 0; q < n; q++) { out[q] = 0; ss >> out[q]; }
  }
};

static inline bool read_raw(const char *path, double *const *val, int nm, long nd) {
  FILE *f = std::fopen(path, "rb");
  if (!f) return false;
  for (int m = 0; m < nm; m++)
    if (std::fread(val[m], sizeof(double), nd, f) != (size_t)nd) { std::fclose(f); return false; }
  std::fclose(f);
  return true;
}

static inline bool write_raw(const char *path, double *const *val, int nm, long nd) {
  FILE *f = std::fopen(path, "wb");
  if (!f) return false;
  for (int m = 0; m < nm; m++) std::fwrite(val[m], sizeof(double), nd, f);
  std::fclose(f);
  return true;
}

static inline bool write_vec(const char *path, const std::vector<double> &v) {
  FILE *f = std::fopen(path, "wb");
  if (!f) return false;
  if (!v.empty()) std::fwrite(v.data(), sizeof(double), v.size(), f);
  std::fclose(f);
  return true;
}
#endif
