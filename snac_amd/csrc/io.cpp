// Checkpoint, visualisation and grid files in the reference's formats
// (io.hpp).  Plain host file I/O: the caller moves the fields between the
// device and the host.
#include "io.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <stdexcept>

namespace snac {
namespace {

struct FileCloser {
  void operator()(std::FILE *f) const {
    if (f) std::fclose(f);
  }
};
using File = std::unique_ptr<std::FILE, FileCloser>;

File open_or_throw(const std::string &path, const char *mode) {
  File f(std::fopen(path.c_str(), mode));
  if (!f) throw std::runtime_error("cannot open " + path);
  return f;
}

void write_doubles(std::FILE *f, const double *v, size_t n, const std::string &path) {
  if (n && std::fwrite(v, sizeof(double), n, f) != n) throw std::runtime_error("short write to " + path);
}

// a real of list-directed output: fixed notation with 17 significant digits
// for moderate magnitudes, exponent notation otherwise
std::string list_real(double x) {
  char b[64];
  const double a = std::fabs(x);
  if (x == 0.0 || (a >= 0.1 && a < 1e16)) {
    const int before = x == 0.0 ? 1 : (int)std::floor(std::log10(a)) + 1;
    std::snprintf(b, sizeof b, "%26.*f", std::max(0, 17 - std::max(before, 1)), x);
  } else {
    std::snprintf(b, sizeof b, "%26.16E", x);
  }
  return b;
}

}  // namespace

std::string fortran_e(double x, int w, int d) {
  std::string body;
  if (std::isnan(x)) {
    body = "NaN";
  } else if (std::isinf(x)) {
    body = x > 0 ? "Infinity" : "-Infinity";
  } else if (x == 0.0) {
    body = std::string(std::signbit(x) ? "-" : "") + "0." + std::string(d, '0') + "E+00";
  } else {
    // d significant digits, rounded to nearest by printf: m.mmm...e+X
    char b[64];
    std::snprintf(b, sizeof b, "%.*e", d - 1, std::fabs(x));
    const std::string s(b);
    const size_t ep = s.find('e');
    const std::string digits = s.substr(0, 1) + (d > 1 ? s.substr(2, ep - 2) : "");
    const int e = std::atoi(s.c_str() + ep + 1) + 1;  // 0.ddd scaling
    char eb[16];
    if (std::abs(e) <= 99) std::snprintf(eb, sizeof eb, "E%c%02d", e < 0 ? '-' : '+', std::abs(e));
    else std::snprintf(eb, sizeof eb, "%c%03d", e < 0 ? '-' : '+', std::abs(e));
    body = std::string(x < 0 ? "-" : "") + "0." + digits + eb;
  }
  if ((int)body.size() > w) return std::string(w, '*');
  return std::string(w - body.size(), ' ') + body;
}

std::string block_num(int b) {
  char s[16];
  std::snprintf(s, sizeof s, "%03d", b + 1);
  return s;
}

void write_field_file(const std::string &path, const double *v, size_t n, double time, int istep) {
  File f = open_or_throw(path, "wb");
  write_doubles(f.get(), v, n, path);
  const double info[2] = {time, 1.0 * istep};
  write_doubles(f.get(), info, 2, path);
}

void read_field_file(const std::string &path, double *v, size_t n, double &time, int &istep) {
  File f = open_or_throw(path, "rb");
  std::fseek(f.get(), 0, SEEK_END);
  const long size = std::ftell(f.get());
  const long good = (long)(n + 2) * (long)sizeof(double);
  if (size != good)
    throw std::runtime_error("*** Simulation aborted due a checkpoint file with incorrect size *** file: " + path +
                             " | expected size: " + std::to_string(good) + " | actual size: " + std::to_string(size));
  std::fseek(f.get(), 0, SEEK_SET);
  double info[2];
  if (std::fread(v, sizeof(double), n, f.get()) != n || std::fread(info, sizeof(double), 2, f.get()) != 2)
    throw std::runtime_error("short read from " + path);
  time = info[0];
  istep = (int)std::lround(info[1]);
}

void write_raw_field(const std::string &path, const double *v, size_t n) {
  File f = open_or_throw(path, "wb");
  write_doubles(f.get(), v, n, path);
}

void append_visu_log(const std::string &path, const std::string &fname, const std::string &varname,
                     const int nmin[3], const int nmax[3], const int nskip[3], double time, int istep) {
  File f = open_or_throw(path, "a");
  std::string line = fname + " " + varname;
  char b[32];
  for (const int *a : {nmin, nmax, nskip})
    for (int q = 0; q < 3; q++) {
      std::snprintf(b, sizeof b, "%5d", a[q]);
      line += (std::string(b).size() > 5) ? std::string(5, '*') : std::string(b);
    }
  line += fortran_e(time, 15, 7);
  std::snprintf(b, sizeof b, "%10d", istep);
  line += b;
  std::fprintf(f.get(), "%s\n", line.c_str());
}

void save_grid(const std::string &base, int n, const std::vector<double> &rf, const std::vector<double> &rc,
               const std::vector<double> &drf, const std::vector<double> &drc) {
  {
    File f = open_or_throw(base + ".bin", "wb");
    for (const std::vector<double> *a : {&rf, &rc, &drf, &drc}) write_doubles(f.get(), a->data() + 1, n, base);
  }
  File f = open_or_throw(base + ".out", "w");
  for (int q = 0; q <= n + 1; q++)
    std::fprintf(f.get(), "%s%s%s%s%s\n", fortran_e(0.0, 15, 7).c_str(), fortran_e(rf[q], 15, 7).c_str(),
                 fortran_e(rc[q], 15, 7).c_str(), fortran_e(drf[q], 15, 7).c_str(), fortran_e(drc[q], 15, 7).c_str());
}

bool load_grid(const std::string &base, int n, std::vector<double> &rf, std::vector<double> &rc,
               std::vector<double> &drf, std::vector<double> &drc) {
  File f(std::fopen((base + ".bin").c_str(), "rb"));
  if (!f) return false;
  for (std::vector<double> *a : {&rf, &rc, &drf, &drc}) {
    a->resize(n + 2);
    if (std::fread(a->data() + 1, sizeof(double), n, f.get()) != (size_t)n)
      throw std::runtime_error("short read from " + base + ".bin");
  }
  drf[0] = drf[1];
  drf[n + 1] = drf[n];
  drc[0] = drf[1];
  drc[n + 1] = drc[n];
  rf[0] = rf[1] - drf[1];
  rf[n + 1] = rf[n] + drf[n];
  rc[0] = rc[1] - drc[0];
  rc[n + 1] = rc[n] + drc[n];
  return true;
}

void write_geometry(const std::string &path, const int lo[3], const int hi[3], const double lmin[3],
                    const double lmax[3]) {
  File f = open_or_throw(path, "w");
  std::fprintf(f.get(), "%12d%12d%12d\n%12d%12d%12d\n", lo[0], lo[1], lo[2], hi[0], hi[1], hi[2]);
  for (const double *a : {lmin, lmax})
    std::fprintf(f.get(), "%s%s%s\n", list_real(a[0]).c_str(), list_real(a[1]).c_str(), list_real(a[2]).c_str());
}

}  // namespace snac
