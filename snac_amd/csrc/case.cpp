// Case set-up: inputs (param.f90:68-277), block placement and neighbours
// (initmpi.f90:9-254, one piece per block since every block lives on one
// GPU), grids (initgrid.f90:9-127), inflow planes (main.f90:600-636 with
// init_inflow initflow.f90:369-429) and pressure BC rhs planes (init_bc_rhs
// solver.f90:44-157).
#include <cmath>
#include <limits>
#include <stdexcept>

#include "case.hpp"
#include "io.hpp"

namespace snac {
namespace {

const NmlValue *find(const NmlGroup &g, const std::string &name, const std::vector<int> &idx) {
  auto it = g.find(name);
  if (it == g.end()) return nullptr;
  auto jt = it->second.find(idx);
  if (jt != it->second.end()) return &jt->second;
  if (idx.empty()) {
    auto kt = it->second.find({1});
    if (kt != it->second.end()) return &kt->second;
  }
  return nullptr;
}
double num(const NmlGroup &g, const std::string &n, double def, const std::vector<int> &idx = {}) {
  auto *v = find(g, n, idx);
  if (!v) return def;
  if (v->kind == NmlValue::Num) return v->num;
  throw std::runtime_error("namelist: " + n + " is not a number");
}
bool logical(const NmlGroup &g, const std::string &n, bool def, const std::vector<int> &idx = {}) {
  auto *v = find(g, n, idx);
  if (!v) return def;
  if (v->kind == NmlValue::Bool) return v->b;
  throw std::runtime_error("namelist: " + n + " is not a logical");
}
std::string str(const NmlGroup &g, const std::string &n, const std::string &def, const std::vector<int> &idx) {
  auto *v = find(g, n, idx);
  if (!v) return def;
  if (v->kind == NmlValue::Str) return v->str;
  throw std::runtime_error("namelist: " + n + " is not a string");
}
char chr(const NmlGroup &g, const std::string &n, char def, const std::vector<int> &idx) {
  auto *v = find(g, n, idx);
  if (!v) return def;
  if (v->kind == NmlValue::Str && !v->str.empty()) return v->str[0];
  throw std::runtime_error("namelist: " + n + " is not a character");
}

double gp_geo_one_end(double a, double r0) {
  const double power = 1.0 / 3.0;
  if (r0 == 1.0) return r0;
  return r0 * std::pow((1.0 - std::pow(r0, a)) / (1.0 - r0) / a, power);
}

// grid stretching functions, initgrid.f90:171-287
double gridpoint(int gt, double a, double r0) {
  switch (gt) {
  case 1: return a != 0.0 ? 1.0 * (1.0 + std::tanh((r0 - 1.0) * a) / std::tanh(a / 1.0)) : r0;
  case 2:
    if (a == 0.0) return r0;
    if (r0 <= 0.5) return 0.5 * (1.0 - 1.0 + std::tanh(2.0 * a * (r0 - 0.0)) / std::tanh(a));
    return 0.5 * (1.0 + 1.0 + std::tanh(2.0 * a * (r0 - 1.0)) / std::tanh(a));
  case 3: return a != 0.0 ? 1.0 - 1.0 * (1.0 + std::tanh((1.0 - r0 - 1.0) * a) / std::tanh(a / 1.0)) : r0;
  case 4: return gp_geo_one_end(a, r0);
  case 5: return 1.0 - gp_geo_one_end(a, 1.0 - r0);
  case 6:
    if (r0 <= 0.5) return gp_geo_one_end(a / 2.0, 2.0 * r0) / 2.0;
    return 1.0 - gp_geo_one_end(a / 2.0, 2.0 * (1.0 - r0)) / 2.0;
  case 7:
    if (r0 <= 0.5) return (1.0 - gp_geo_one_end(a / 2.0, (0.5 - r0) * 2.0)) / 2.0;
    return (1.0 + gp_geo_one_end(a / 2.0, (r0 - 0.5) * 2.0)) / 2.0;
  case 0:
  default: return a != 0.0 ? 0.5 * (1.0 + std::tanh((r0 - 0.5) * a) / std::tanh(a / 2.0)) : r0;
  }
}

void tangential(int d, int &a1, int &a2) {
  a1 = d == 0 ? 1 : 0;
  a2 = d == 2 ? 1 : 2;
}

}  // namespace

void initgrid(int n, int gt, double gr, double lmin, double lmax, std::vector<double> &drc,
              std::vector<double> &drf, std::vector<double> &rc, std::vector<double> &rf) {
  drc.assign(n + 2, 0.0); drf.assign(n + 2, 0.0); rc.assign(n + 2, 0.0); rf.assign(n + 2, 0.0);
  for (int q = 1; q <= n; q++) rf[q] = lmin + gridpoint(gt, gr, (double)q / (double)n) * (lmax - lmin);
  rf[0] = lmin;
  for (int q = 1; q <= n; q++) drf[q] = rf[q] - rf[q - 1];
  drf[0] = drf[1];
  drf[n + 1] = drf[n];
  for (int q = 0; q <= n; q++) drc[q] = 0.5 * (drf[q] + drf[q + 1]);
  drc[n + 1] = drc[n];
  rc[0] = lmin - drc[0] / 2.0;
  rf[0] = lmin;
  for (int q = 1; q <= n + 1; q++) {
    rc[q] = rc[q - 1] + drc[q - 1];
    rf[q] = rf[q - 1] + drf[q];
  }
}

bool find_fft(char a, char b, int &kf, int &kb, double norm[2], bool centred) {
  // FFTW kind numbers (fftw.f90:31-54)
  if (a == 'F' && b == 'F') { kf = 0; kb = 1; norm[0] = 1; norm[1] = 0; return true; }
  if (!centred) {  // staggered: DD only (RODFT00 over n-1 points, norm 2(n+1-1))
    if (a == 'D' && b == 'D') { kf = 7; kb = 7; norm[0] = 2; norm[1] = 1; return true; }
    return false;
  }
  if (a == 'N' && b == 'N') { kf = 5; kb = 4; norm[0] = 2; norm[1] = 0; return true; }
  if (a == 'D' && b == 'D') { kf = 9; kb = 8; norm[0] = 2; norm[1] = 0; return true; }
  if (a == 'N' && b == 'D') { kf = 6; kb = 6; norm[0] = 2; norm[1] = 0; return true; }
  if (a == 'D' && b == 'N') { kf = 10; kb = 10; norm[0] = 2; norm[1] = 0; return true; }
  return false;
}

bool eigenvalues(int n, char a, char b, std::vector<double> &lam, bool centred) {
  const double pi = std::acos(-1.0);
  lam.assign(n, 0.0);
  if (!centred && !(a == 'F' && b == 'F')) {
    if (!(a == 'D' && b == 'D')) return false;
    for (int l = 1; l <= n - 1; l++) {
      const double s = std::sin((double)l * pi / (2.0 * n));
      lam[l - 1] = -4.0 * s * s;
    }
    return true;
  }
  for (int l = 1; l <= n; l++) {
    double s;
    if (a == 'F' && b == 'F') s = std::sin((double)(l - 1) * pi / (double)n);
    else if (a == 'N' && b == 'N') s = std::sin((double)(l - 1) * pi / (2.0 * n));
    else if (a == 'D' && b == 'D') s = std::sin((double)l * pi / (2.0 * n));
    else if ((a == 'N' && b == 'D') || (a == 'D' && b == 'N')) s = std::sin((double)(2 * l - 1) * pi / (4.0 * n));
    else return false;
    lam[l - 1] = -4.0 * s * s;
  }
  return true;
}

Case build_case(const std::string &dns_path, const std::string &blocks_path, int fft_axis) {
  if (fft_axis < 0 || fft_axis > 3) throw std::runtime_error("fft_axis must be 0..3");
  Nml dnsf = read_namelist_file(dns_path);
  Nml blkf = read_namelist_file(blocks_path);
  if (!dnsf.count("dns")) throw std::runtime_error("Error reading dns namelist");
  if (!blkf.count("blocks")) throw std::runtime_error("Error reading blocks namelist");
  const NmlGroup &d = dnsf["dns"];
  const NmlGroup &bl = blkf["blocks"];
  Case c;
  c.fft_axis = fft_axis;
  c.cfl = num(d, "cfl", 0.0);
  c.dtmax = num(d, "dtmax", 0.0);
  c.dt_f = num(d, "dt_f", -1.0);
  c.visci = num(d, "visci", 1.0);
  c.visc = 1.0 / c.visci;
  c.nstep = (int)num(d, "nstep", 0);
  c.time_max = num(d, "time_max", 0.0);
  c.tw_max = num(d, "tw_max", 0.0);
  for (int q = 0; q < 3; q++) c.stop_type[q] = logical(d, "stop_type", false, {q + 1});
  c.restart = logical(d, "restart", false);
  c.is_overwrite_save = logical(d, "is_overwrite_save", true);
  c.nsaves_max = (int)num(d, "nsaves_max", 0);
  c.icheck = (int)num(d, "icheck", 0);
  c.iout0d = (int)num(d, "iout0d", 0);
  c.iout1d = (int)num(d, "iout1d", 0);
  c.iout2d = (int)num(d, "iout2d", 0);
  c.iout3d = (int)num(d, "iout3d", 0);
  c.isave = (int)num(d, "isave", 0);
  for (int q = 0; q < 3; q++) {
    c.bforce[q] = num(d, "bforce", 0.0, {q + 1});
    c.gacc[q] = num(d, "gacc", 0.0, {q + 1});
  }
  c.nscal = (int)num(d, "nscal", 0);
  if (c.nscal < 0) c.nscal = 0;
  if (c.nscal > 8) throw std::runtime_error("at most 8 scalars are supported");
  if (c.nscal > 0) {
    if (!dnsf.count("scalar")) throw std::runtime_error("Error reading scalar namelist");
    const NmlGroup &sg = dnsf["scalar"];
    for (int i = 0; i < c.nscal; i++) {
      const double ai = num(sg, "alphai", std::numeric_limits<double>::max(), {i + 1});
      c.alpha.push_back(1.0 / ai);
      c.ssource.push_back(num(sg, "ssource", 0.0, {i + 1}));
      c.is_sforced.push_back(logical(sg, "is_sforced", false, {i + 1}) ? 1 : 0);
      c.scalf.push_back(num(sg, "scalf", 0.0, {i + 1}));
      c.iniscal.push_back(str(sg, "iniscal", "zer", {i + 1}));
    }
    c.beta = num(sg, "beta", 0.0);
    c.alpha_max = *std::max_element(c.alpha.begin(), c.alpha.end());
  }
  if (dnsf.count("hypre")) {
    const NmlGroup &h = dnsf["hypre"];
    c.hypre_solver_i = (int)num(h, "hypre_solver_i", 2);
    c.hypre_tol = num(h, "hypre_tol", 1e-4);
    c.hypre_maxiter = (int)num(h, "hypre_maxiter", 50);
  }
  int nb = (int)num(bl, "nblocks", 0);
  if (nb < 1 || nb > 999) throw std::runtime_error("invalid number of blocks in blocks.nml");
  c.blocks.resize(nb);
  for (int b = 0; b < nb; b++) {
    Block &B = c.blocks[b];
    for (int q = 0; q < 3; q++) {
      B.n[q] = (int)num(bl, "block_ng", 1, {q + 1, b + 1});
      B.lmin[q] = num(bl, "block_lmin", 0.0, {q + 1, b + 1});
      B.lmax[q] = num(bl, "block_lmax", 1.0, {q + 1, b + 1});
      B.gt[q] = (int)num(bl, "block_gt", 0, {q + 1, b + 1});
      B.gr[q] = num(bl, "block_gr", 0.0, {q + 1, b + 1});
      if (q == 0) B.inivel = str(bl, "block_inivel", "zer", {b + 1}).substr(0, 3);
      if (B.n[q] < 1) throw std::runtime_error("invalid block_ng");
      if (B.gr[q] < 0) throw std::runtime_error("grid growth parameter must be positive.");
      B.cbcsca.resize((size_t)c.nscal * 6, 'N');
      B.bcsca.resize((size_t)c.nscal * 6, 0.0);
      for (int s = 0; s < 2; s++) {
        for (int i = 0; i < c.nscal; i++) {
          B.cbcsca[(size_t)i * 6 + q * 2 + s] = chr(bl, "block_cbcscal", 'N', {s, q + 1, i + 1, b + 1});
          B.bcsca[(size_t)i * 6 + q * 2 + s] = num(bl, "block_bcscal", 0.0, {s, q + 1, i + 1, b + 1});
        }
        B.cbcpre[q][s] = chr(bl, "block_cbcpre", 'N', {s, q + 1, b + 1});
        B.bcpre[q][s] = num(bl, "block_bcpre", 0.0, {s, q + 1, b + 1});
        B.inflow_type[q][s] = (int)num(bl, "block_inflow_type", 0, {s, q + 1, b + 1});
        for (int cp = 0; cp < 3; cp++) {
          B.cbcvel[cp][q][s] = chr(bl, "block_cbcvel", 'D', {s, q + 1, cp + 1, b + 1});
          B.bcvel[cp][q][s] = num(bl, "block_bcvel", 0.0, {s, q + 1, cp + 1, b + 1});
        }
      }
    }
  }
  // chk_bc sanity.f90:103-161 (validity of each pair)
  auto valid_pair = [](char a, char b) {
    std::string p{a, b};
    for (const char *v : {"ND", "DN", "NN", "DD", "FD", "DF", "FF", "FN", "NF"})
      if (p == v) return true;
    return false;
  };
  for (auto &B : c.blocks)
    for (int q = 0; q < 3; q++) {
      if (!valid_pair(B.cbcpre[q][0], B.cbcpre[q][1])) throw std::runtime_error("pressure BCs not valid.");
      for (int cp = 0; cp < 3; cp++)
        if (!valid_pair(B.cbcvel[cp][q][0], B.cbcvel[cp][q][1])) throw std::runtime_error("velocity BCs not valid.");
    }
  // chk_bc sanity.f90:136-157: the normal velocity component's pair and the
  // pressure pair of each direction must be one of these combinations
  for (auto &B : c.blocks)
    for (int q = 0; q < 3; q++) {
      const std::string v{B.cbcvel[q][q][0], B.cbcvel[q][q][1]}, p{B.cbcpre[q][0], B.cbcpre[q][1]};
      static const char *ok[][2] = {{"FF", "FF"}, {"ND", "DN"}, {"DN", "ND"}, {"DD", "NN"}, {"FD", "FN"},
                                    {"DF", "NF"}, {"FN", "FD"}, {"NF", "DF"}, {"FN", "FN"}, {"NF", "NF"},
                                    {"NN", "NN"}, {"DN", "NN"}, {"ND", "NN"}, {"NN", "DD"}};
      bool found = false;
      for (auto &o : ok) found = found || (v == o[0] && p == o[1]);
      if (!found) throw std::runtime_error("velocity and pressure BCs not compatible.");
    }
  // chk_stop_type sanity.f90:41-51
  if (!c.stop_type[0] && !c.stop_type[1] && !c.stop_type[2]) throw std::runtime_error("stopping criterion not chosen.");
  // is_periodic param.f90:220-225
  for (int q = 0; q < 3; q++) {
    c.is_periodic[q] = true;
    for (auto &B : c.blocks) c.is_periodic[q] = c.is_periodic[q] && B.cbcpre[q][0] == 'F' && B.cbcpre[q][1] == 'F';
  }
  // block placement initmpi.f90:61-98: block 1 at lo = 1, breadth-first over
  // 'F' links; links to blocks at a lower (higher) lmin through the upper
  // (lower) side are periodic wraps and are skipped
  {
    std::vector<bool> done(nb, false), unl(nb, false);
    c.blocks[0].lo_g = {1, 1, 1};
    unl[0] = true;
    int ib = 0;
    for (;;) {
      Block &B = c.blocks[ib];
      for (int q = 0; q < 3; q++)
        for (int s = 0; s < 2; s++) {
          if (B.cbcpre[q][s] != 'F') continue;
          int fr = (int)std::lround(B.bcpre[q][s]) - 1;
          if (fr < 0 || fr >= nb) throw std::runtime_error("invalid connectivity for block " + std::to_string(ib + 1));
          if (unl[fr]) continue;
          Block &F = c.blocks[fr];
          if (s == 0 && F.lmin[q] > B.lmin[q]) continue;
          if (s == 1 && F.lmin[q] < B.lmin[q]) continue;
          int off = -(1 - s) * F.n[q] + s * B.n[q];
          F.lo_g = B.lo_g;
          F.lo_g[q] = B.lo_g[q] + off;
          unl[fr] = true;
        }
      done[ib] = true;
      int next = -1;
      bool all = true;
      for (int q = 0; q < nb; q++) {
        all = all && done[q];
        if (next < 0 && !done[q] && unl[q]) next = q;
      }
      if (all) break;
      if (next < 0) throw std::runtime_error("invalid connectivity for block");
      ib = next;
    }
  }
  // periods initmpi.f90:104-113
  for (int q = 0; q < 3; q++) {
    int mn = 1 << 30, mx = -(1 << 30);
    for (auto &B : c.blocks) {
      mn = std::min(mn, B.lo_g[q]);
      mx = std::max(mx, B.lo_g[q] + B.n[q] - 1);
    }
    c.periods[q] = c.is_periodic[q] ? mx - mn + 1 : 0;
  }
  // neighbours initmpi.f90:150-239
  for (int b = 0; b < nb; b++) {
    Block &B = c.blocks[b];
    for (int q = 0; q < 3; q++)
      for (int s = 0; s < 2; s++) {
        B.nb[q][s] = -1;
        if (B.cbcpre[q][s] != 'F') continue;
        int fr = (int)std::lround(B.bcpre[q][s]) - 1;
        const Block &F = c.blocks[fr];
        bool ok = F.cbcpre[q][1 - s] == 'F' && (int)std::lround(F.bcpre[q][1 - s]) - 1 == b;
        for (int e = 0; e < 3; e++) {
          if (e == q) continue;
          ok = ok && F.lo_g[e] == B.lo_g[e] && F.n[e] == B.n[e] && F.lmin[e] == B.lmin[e] &&
               F.lmax[e] == B.lmax[e] && F.gr[e] == B.gr[e] && F.gt[e] == B.gt[e];
        }
        int blo = B.lo_g[q], bhi = B.lo_g[q] + B.n[q] - 1, flo = F.lo_g[q], fhi = F.lo_g[q] + F.n[q] - 1;
        if (s == 0) ok = ok && (blo == fhi + 1 || c.periods[q] == fhi - blo + 1);
        else ok = ok && (bhi == flo - 1 || c.periods[q] == bhi - flo + 1);
        if (!ok)
          throw std::runtime_error("Expected connectivity between blocks " + std::to_string(b + 1) + " and " +
                                   std::to_string(fr + 1) + " is not possible.");
        B.nb[q][s] = fr;
      }
    for (int q = 0; q < 3; q++)
      for (int s = 0; s < 2; s++) B.is_bound[q][s] = B.nb[q][s] < 0;
  }
  // chk_grid_fft sanity.f90:66-100
  if (fft_axis > 0) {
    int a = fft_axis - 1;
    for (auto &B : c.blocks) {
      if (B.gr[a] != 0.0) throw std::runtime_error("grid growth parameter in the FFT direction must be zero.");
      if (B.n[a] != c.blocks[0].n[a] || B.lmin[a] != c.blocks[0].lmin[a] || B.lmax[a] != c.blocks[0].lmax[a])
        throw std::runtime_error("domain size and number of points cannot vary among blocks in the FFT direction");
      if (B.cbcpre[a][0] != c.blocks[0].cbcpre[a][0] || B.cbcpre[a][1] != c.blocks[0].cbcpre[a][1])
        throw std::runtime_error("pressure BCs along the FFT direction must agree among blocks");
    }
  }
  // vol_all param.f90:239-240
  c.vol_all = 0;
  for (auto &B : c.blocks) c.vol_all += (B.lmax[0] - B.lmin[0]) * (B.lmax[1] - B.lmin[1]) * (B.lmax[2] - B.lmin[2]);
  // grids: initgrid then bound_grid (initgrid.f90:84-127), the exchanged
  // ghost spacing being the neighbour's interior spacing
  // a grid file grid/grid_{x,y,z}_b_NNN.bin in the working directory
  // replaces the generated grid of that block and direction (main.f90:415-428)
  for (size_t b = 0; b < c.blocks.size(); b++) {
    Block &B = c.blocks[b];
    for (int q = 0; q < 3; q++) {
      initgrid(B.n[q], B.gt[q], B.gr[q], B.lmin[q], B.lmax[q], B.dc[q], B.df[q], B.rc[q], B.rf[q]);
      const std::string base = std::string("grid/grid_") + "xyz"[q] + "_b_" + block_num((int)b);
      load_grid(base, B.n[q], B.rf[q], B.rc[q], B.df[q], B.dc[q]);
      B.df0[q] = B.df[q];
      B.dc0[q] = B.dc[q];
    }
  }
  {
    std::vector<std::array<double, 3>> lov(nb), hiv(nb);
    for (int b = 0; b < nb; b++)
      for (int q = 0; q < 3; q++) { lov[b][q] = c.blocks[b].df[q][1]; hiv[b][q] = c.blocks[b].df[q][c.blocks[b].n[q]]; }
    for (auto &B : c.blocks)
      for (int q = 0; q < 3; q++) {
        int n = B.n[q];
        if (B.nb[q][0] >= 0) B.df[q][0] = hiv[B.nb[q][0]][q];
        if (B.nb[q][1] >= 0) B.df[q][n + 1] = lov[B.nb[q][1]][q];
        B.dc[q][0] = (B.df[q][1] + B.df[q][0]) / 2.0;
        B.dc[q][n] = (B.df[q][n + 1] + B.df[q][n]) / 2.0;
        B.dc[q][n + 1] = B.dc[q][n];
      }
  }
  // chk_inflow sanity.f90:181-198, then inflow planes
  for (auto &B : c.blocks)
    for (int q = 0; q < 3; q++)
      for (int s = 0; s < 2; s++) {
        if (B.inflow_type[q][s] > 0 && c.periods[q] != 0)
          throw std::runtime_error("Periodicity and inflow BCs cannot be combined.");
        B.is_inflow[q][s] = B.is_bound[q][s] && B.inflow_type[q][s] > 0 && B.cbcvel[q][q][s] == 'D';
        if (!B.is_inflow[q][s]) continue;
        int a1, a2;
        tangential(q, a1, a2);
        int m1 = B.n[a1] + 2, m2 = B.n[a2] + 2;
        for (int cp = 0; cp < 3; cp++) B.vin[q][cp][s].assign((size_t)m1 * m2, 0.0);
        double uref = B.bcvel[q][q][s];
        bool q1 = c.periods[a1] == 0, q2 = c.periods[a2] == 0;
        auto pois = [](double r, double lo, double hi) {
          double rr = (r - lo) / (hi - lo);
          return 6.0 * (rr * (1.0 - rr));
        };
        if (B.inflow_type[q][s] == 1) {
          for (int t2 = 0; t2 < m2; t2++)
            for (int t1 = 0; t1 < m1; t1++) {
              double f1 = q1 ? pois(B.rc[a1][t1], B.lmin[a1], B.lmax[a1]) : 1.0;
              double f2 = q2 ? pois(B.rc[a2][t2], B.lmin[a2], B.lmax[a2]) : 1.0;
              B.vin[q][q][s][t1 + (size_t)m1 * t2] = uref * f1 * f2;
            }
        } else if (B.inflow_type[q][s] == 2) {
          // Kovasznay flow at x0 = -0.5 with lref = 1 (initflow.f90:407-416):
          // normal 1 - e^(lambda x0) cos(2 pi x2c), first tangential 0,
          // second tangential lambda/(2 pi) e^(lambda x0) sin(2 pi x2f), x2 the
          // second tangential axis
          const double pi = std::acos(-1.0);
          const double reb = uref * 1.0 / c.visc;
          const double lambda = reb / 2 - std::sqrt(reb * reb / 4 + 4 * (pi * pi));
          const double x0 = -0.5;
          for (int t2 = 0; t2 < m2; t2++)
            for (int t1 = 0; t1 < m1; t1++) {
              const size_t e = t1 + (size_t)m1 * t2;
              B.vin[q][q][s][e] = 1.0 - std::exp(lambda * x0) * std::cos(2 * pi * B.rc[a2][t2]);
              B.vin[q][a2][s][e] = lambda / (2.0 * pi) * std::exp(lambda * x0) * std::sin(2 * pi * B.rf[a2][t2]);
            }
        } else if (B.inflow_type[q][s] != 0) {
          throw std::runtime_error("inflow type " + std::to_string(B.inflow_type[q][s]) + " is not implemented");
        }
      }
  // pressure BC rhs planes, init_bc_rhs solver.f90:44-157 (centered);
  // dl = dxc_g(lo_g-1), dxc_g(hi_g) = interior face spacings (main.f90:717-719)
  for (auto &B : c.blocks)
    for (int q = 0; q < 3; q++) {
      if (c.periods[q] != 0) continue;
      int a1, a2;
      tangential(q, a1, a2);
      for (int s = 0; s < 2; s++) {
        if (!B.is_bound[q][s]) continue;
        double dl = s == 0 ? B.df[q][1] : B.df[q][B.n[q]];
        double factor = 0;
        if (B.cbcpre[q][s] == 'N') factor = (s == 1 ? -1.0 : 1.0) * dl * B.bcpre[q][s];
        else if (B.cbcpre[q][s] == 'D') factor = -2.0 * B.bcpre[q][s];
        int qq = s == 0 ? 1 : B.n[q];
        int m1 = B.n[a1], m2 = B.n[a2];
        B.rhsp[q][s].assign((size_t)m1 * m2, 0.0);
        for (int t2 = 1; t2 <= m2; t2++)
          for (int t1 = 1; t1 <= m1; t1++) {
            int ijk[3];
            ijk[q] = qq; ijk[a1] = t1; ijk[a2] = t2;
            double w;
            switch (fft_axis) {
            case 1: w = B.df[1][ijk[1]] * B.df[2][ijk[2]]; break;
            case 2: w = B.df[0][ijk[0]] * B.df[2][ijk[2]]; break;
            case 3: w = B.df[0][ijk[0]] * B.df[1][ijk[1]]; break;
            default: w = B.df[0][ijk[0]] * B.df[1][ijk[1]] * B.df[2][ijk[2]]; break;
            }
            B.rhsp[q][s][(t1 - 1) + (size_t)m1 * (t2 - 1)] = factor * w / (B.dc[q][qq - (1 - s)] * B.df[q][qq]);
          }
      }
    }
  return c;
}

}  // namespace snac
