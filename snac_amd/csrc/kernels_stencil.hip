// Volume kernels of one RK3 substage: the fused momentum RHS + low-storage
// RK3 update (rk.f90:7-128, mom.f90:112-386), fillps (fillps.f90:7-54) with
// updt_rhs (bound.f90:543-585), correc (correc.f90:7-48) with updt_pressure
// (updt_pressure.f90:23-29), and the chkdiv / chkdt reductions
// (chkdiv.f90:8-44, chkdt.f90:8-56).  FP64 throughout; one thread per cell,
// 64-wide rows along i so that every wavefront reads whole 512-byte rows.
#include "common.hpp"

namespace snac {
namespace {

constexpr int TX = 64, TY = 4;

__device__ __forceinline__ long gidx(long i, long j, long k, long sx, long sy) { return i + sx * (j + sy * k); }

// Tile coordinates of this workgroup in a (gx, gy, *) tile grid launched as a
// 1D grid.  Workgroups are dispatched round-robin over the 8 XCDs; remapping
// gives each XCD one contiguous run of tiles (a k-slab), so the i/j/k
// neighbours a stencil re-reads sit in the same XCD's L2.  A bijection for
// any grid size.
__device__ __forceinline__ void tile_of(int gx, int gy, int &bx, int &by, int &bz) {
  const int nblk = gridDim.x, bid = blockIdx.x;
  const int xcd = bid & 7, q = bid >> 3, per = nblk >> 3, rem = nblk & 7;
  const int lid = (xcd < rem ? xcd * (per + 1) : rem * (per + 1) + (xcd - rem) * per) + q;
  bx = lid % gx;
  const int t = lid / gx;
  by = t % gy;
  bz = t / gy;
}

// The old tendencies dudtrko, dvdtrko, dwdtrko share one array, row by row:
// [k][j][component][i], so the three streams of a tile row sit side by side
// in HBM (dtv = dtu + nx, dtw = dtu + 2 nx); this is the index of interior
// cell (i, j, k) (1-based) in component 0.
__device__ __forceinline__ long dto_index(const Geo &g, int i, int j, int k) {
  return (long)(i - 1) + 3L * g.nx * ((long)(j - 1) + (long)g.ny * (long)(k - 1));
}

// BUOY: Boussinesq buoyancy of scalar 1, -f12*gacc*beta*(face average of s)
// (rk.f90:104-108), gb = gacc*beta
template <bool BUOY, bool IMP>
__global__ __launch_bounds__(TX *TY) void k_rk_mom(Geo g, Metrics m, const double *__restrict__ u,
                                                    const double *__restrict__ v, const double *__restrict__ w,
                                                    const double *__restrict__ p, double *__restrict__ un,
                                                    double *__restrict__ vn, double *__restrict__ wn,
                                                    double *__restrict__ dtu, double *__restrict__ dtv,
                                                    double *__restrict__ dtw, double f1, double f2, double f12,
                                                    double visc, double bfx, double bfy, double bfz,
                                                    const double *__restrict__ s1, double gbx, double gby,
                                                    double gbz) {
  int bx, by, bz;
  tile_of((g.nx + TX - 1) / TX, (g.ny + TY - 1) / TY, bx, by, bz);
  const int i = bx * TX + threadIdx.x + 1;
  const int j = by * TY + threadIdx.y + 1;
  const int k = bz + 1;
  if (i > g.nx || j > g.ny) return;
  const long sx = g.sx, sy = g.sy;
  const long c = gidx(i, j, k, sx, sy);
  const long xp = c + 1, xm = c - 1, yp = c + sx, ym = c - sx, zp = c + sx * sy, zm = c - sx * sy;
  const double dxfi = m.dxf[i], dxfip = m.dxf[i + 1], dyfj = m.dyf[j], dyfjp = m.dyf[j + 1];
  const double dzfk = m.dzf[k], dzfkp = m.dzf[k + 1];
  const double ixci = m.idxc[i], ixcim = m.idxc[i - 1], ixfi = m.idxf[i], ixfip = m.idxf[i + 1];
  const double iycj = m.idyc[j], iycjm = m.idyc[j - 1], iyfj = m.idyf[j], iyfjp = m.idyf[j + 1];
  const double izck = m.idzc[k], izckm = m.idzc[k - 1], izfk = m.idzf[k], izfkp = m.idzf[k + 1];
  const double uc = u[c], vc = v[c], wc = w[c];
  double a, b, cc;
  // diffusion, momx_d / momy_d / momz_d (mom.f90:221-327)
  {
    double dudxp = (u[xp] - uc) * ixfip, dudxm = (uc - u[xm]) * ixfi;
    double dudyp = (u[yp] - uc) * iycj, dudym = (uc - u[ym]) * iycjm;
    double dudzp = (u[zp] - uc) * izck, dudzm = (uc - u[zm]) * izckm;
    a = visc * (dudxp - dudxm) * ixci + visc * (dudyp - dudym) * iyfj + visc * (dudzp - dudzm) * izfk;
    double dvdxp = (v[xp] - vc) * ixci, dvdxm = (vc - v[xm]) * ixcim;
    double dvdyp = (v[yp] - vc) * iyfjp, dvdym = (vc - v[ym]) * iyfj;
    double dvdzp = (v[zp] - vc) * izck, dvdzm = (vc - v[zm]) * izckm;
    b = visc * (dvdxp - dvdxm) * ixfi + visc * (dvdyp - dvdym) * iycj + visc * (dvdzp - dvdzm) * izfk;
    double dwdxp = (w[xp] - wc) * ixci, dwdxm = (wc - w[xm]) * ixcim;
    double dwdyp = (w[yp] - wc) * iycj, dwdym = (wc - w[ym]) * iycjm;
    double dwdzp = (w[zp] - wc) * izfkp, dwdzm = (wc - w[zm]) * izfk;
    cc = visc * (dwdxp - dwdxm) * ixfi + visc * (dwdyp - dwdym) * iyfj + visc * (dwdzp - dwdzm) * izck;
  }
  // implicit diffusion: the diffusion is kept apart, dudtrk is the advection
  // alone (rk.f90:50-59)
  double ad = 0.0, bd = 0.0, cd = 0.0;
  if (IMP) { ad = a; bd = b; cd = cc; a = 0.0; b = 0.0; cc = 0.0; }
  // volume-weighted advection, momx/y/z_a_vv (mom.f90:112-219)
  {
    double uuip = 0.25 * (uc * dyfj * dzfk + u[xp] * dyfj * dzfk) * u[xp];
    double uuim = 0.25 * (uc * dyfj * dzfk + u[xm] * dyfj * dzfk) * u[xm];
    double uvjp = 0.25 * (vc * dxfi * dzfk + v[xp] * dxfip * dzfk) * u[yp];
    double uvjm = 0.25 * (v[ym] * dxfi * dzfk + v[ym + 1] * dxfip * dzfk) * u[ym];
    double uwkp = 0.25 * (wc * dxfi * dyfj + w[xp] * dxfip * dyfj) * u[zp];
    double uwkm = 0.25 * (w[zm] * dxfi * dyfj + w[zm + 1] * dxfip * dyfj) * u[zm];
    a = a + ((-uuip + uuim) + (-uvjp + uvjm) + (-uwkp + uwkm)) * (ixci * iyfj * izfk);
    double uvip = 0.25 * (uc * dyfj * dzfk + u[yp] * dyfjp * dzfk) * v[xp];
    double uvim = 0.25 * (u[xm] * dyfj * dzfk + u[xm + sx] * dyfjp * dzfk) * v[xm];
    double vvjp = 0.25 * (vc * dxfi * dzfk + v[yp] * dxfi * dzfk) * v[yp];
    double vvjm = 0.25 * (vc * dxfi * dzfk + v[ym] * dxfi * dzfk) * v[ym];
    double wvkp = 0.25 * (wc * dyfj * dxfi + w[yp] * dyfjp * dxfi) * v[zp];
    double wvkm = 0.25 * (w[zm] * dyfj * dxfi + w[zm + sx] * dyfjp * dxfi) * v[zm];
    b = b + ((-uvip + uvim) + (-vvjp + vvjm) + (-wvkp + wvkm)) * (ixfi * iycj * izfk);
    double uwip = 0.25 * (uc * dzfk * dyfj + u[zp] * dzfkp * dyfj) * w[xp];
    double uwim = 0.25 * (u[xm] * dzfk * dyfj + u[xm + sx * sy] * dzfkp * dyfj) * w[xm];
    double vwjp = 0.25 * (vc * dzfk * dxfi + v[zp] * dzfkp * dxfi) * w[yp];
    double vwjm = 0.25 * (v[ym] * dzfk * dxfi + v[ym + sx * sy] * dzfkp * dxfi) * w[ym];
    double wwkp = 0.25 * (wc * dxfi * dyfj + w[zp] * dxfi * dyfj) * w[zp];
    double wwkm = 0.25 * (wc * dxfi * dyfj + w[zm] * dxfi * dyfj) * w[zm];
    cc = cc + ((-uwip + uwim) + (-vwjp + vwjm) + (-wwkp + wwkm)) * (ixfi * iyfj * izck);
  }
  // RK3 update (rk.f90:68-85), then pressure gradient + body force (86-111)
  const long q = dto_index(g, i, j, k);
  double un_ = uc + f1 * a + f2 * dtu[q];
  double vn_ = vc + f1 * b + f2 * dtv[q];
  double wn_ = wc + f1 * cc + f2 * dtw[q];
  if (IMP) { un_ = un_ + f12 * ad; vn_ = vn_ + f12 * bd; wn_ = wn_ + f12 * cd; }  // rk.f90:75-79
  dtu[q] = a;
  dtv[q] = b;
  dtw[q] = cc;
  const double pc = p[c];
  un_ = un_ + f12 * (-(p[xp] - pc) * ixci + bfx);
  vn_ = vn_ + f12 * (-(p[yp] - pc) * iycj + bfy);
  wn_ = wn_ + f12 * (-(p[zp] - pc) * izck + bfz);
  if (BUOY) {
    const double sc = s1[c];
    un_ = un_ - f12 * gbx * 0.5 * (s1[xp] + sc);
    vn_ = vn_ - f12 * gby * 0.5 * (s1[yp] + sc);
    wn_ = wn_ - f12 * gbz * 0.5 * (s1[zp] + sc);
  }
  if (IMP) {  // the Helmholtz right-hand side (rk.f90:112-127)
    un_ = un_ - 0.5 * f12 * ad;
    vn_ = vn_ - 0.5 * f12 * bd;
    wn_ = wn_ - 0.5 * f12 * cd;
  }
  un[c] = un_;
  vn[c] = vn_;
  wn[c] = wn_;
}

// The same substage with each thread marching along k over a chunk of KB
// planes: its own column's values at k-1, k, k+1 of u, v, w and at k, k+1
// of p stay in registers from one plane to the next, so every plane of the
// fields is fetched once per column; the in-plane neighbours and the few
// mixed (i-1,k+1)/(j-1,k+1)/(i+1,k-1)/(j+1,k-1) terms come through L1/L2
// from loads that neighbouring lanes and wavefronts issue for their own
// columns.  Metrics along i, j are hoisted out of the march; those along k
// are uniform per plane.
// BX x BY threads per workgroup (256).  (Capping the registers for more
// waves per SIMD measured slower: the per-wave load parallelism matters
// more than the wave count here.)
template <bool BUOY, bool IMP, int KB, int BX, int BY>
__global__ __launch_bounds__(BX *BY) void k_rk_mom_kz(Geo g, Metrics m, const double *__restrict__ u,
                                                       const double *__restrict__ v, const double *__restrict__ w,
                                                       const double *__restrict__ p, double *__restrict__ un,
                                                       double *__restrict__ vn, double *__restrict__ wn,
                                                       double *__restrict__ dtu, double *__restrict__ dtv,
                                                       double *__restrict__ dtw, double f1, double f2, double f12,
                                                       double visc, double bfx, double bfy, double bfz,
                                                       const double *__restrict__ s1, double gbx, double gby,
                                                       double gbz) {
  int bx, by, bz;
  tile_of((g.nx + BX - 1) / BX, (g.ny + BY - 1) / BY, bx, by, bz);
  const int i = bx * BX + threadIdx.x + 1;
  const int j = by * BY + threadIdx.y + 1;
  const int k0 = bz * KB + 1;
  const int k1 = min(k0 + KB - 1, g.nz);
  if (i > g.nx || j > g.ny) return;
  const long sx = g.sx, sy = g.sy, sz = sx * sy;
  const double dxfi = m.dxf[i], dxfip = m.dxf[i + 1], dyfj = m.dyf[j], dyfjp = m.dyf[j + 1];
  const double ixci = m.idxc[i], ixcim = m.idxc[i - 1], ixfi = m.idxf[i], ixfip = m.idxf[i + 1];
  const double iycj = m.idyc[j], iycjm = m.idyc[j - 1], iyfj = m.idyf[j], iyfjp = m.idyf[j + 1];
  long c = gidx(i, j, k0, sx, sy);
  double um = u[c - sz], vm = v[c - sz], wm = w[c - sz];
  double uc = u[c], vc = v[c], wc = w[c], pc = p[c];
  long q = dto_index(g, i, j, k0);
  const long qs = 3L * g.nx * g.ny;
  for (int k = k0; k <= k1; k++, c += sz, q += qs) {
    const long xp = c + 1, xm = c - 1, yp = c + sx, ym = c - sx, zp = c + sz, zm = c - sz;
    const double up = u[zp], vp = v[zp], wp = w[zp], pz = p[zp];
    const double dzfk = m.dzf[k], dzfkp = m.dzf[k + 1];
    const double izck = m.idzc[k], izckm = m.idzc[k - 1], izfk = m.idzf[k], izfkp = m.idzf[k + 1];
    const double uxp = u[xp], uxm = u[xm], uyp = u[yp], uym = u[ym];
    const double vxp = v[xp], vxm = v[xm], vyp = v[yp], vym = v[ym];
    const double wxp = w[xp], wxm = w[xm], wyp = w[yp], wym = w[ym];
    double a, b, cc;
    {
      double dudxp = (uxp - uc) * ixfip, dudxm = (uc - uxm) * ixfi;
      double dudyp = (uyp - uc) * iycj, dudym = (uc - uym) * iycjm;
      double dudzp = (up - uc) * izck, dudzm = (uc - um) * izckm;
      a = visc * (dudxp - dudxm) * ixci + visc * (dudyp - dudym) * iyfj + visc * (dudzp - dudzm) * izfk;
      double dvdxp = (vxp - vc) * ixci, dvdxm = (vc - vxm) * ixcim;
      double dvdyp = (vyp - vc) * iyfjp, dvdym = (vc - vym) * iyfj;
      double dvdzp = (vp - vc) * izck, dvdzm = (vc - vm) * izckm;
      b = visc * (dvdxp - dvdxm) * ixfi + visc * (dvdyp - dvdym) * iycj + visc * (dvdzp - dvdzm) * izfk;
      double dwdxp = (wxp - wc) * ixci, dwdxm = (wc - wxm) * ixcim;
      double dwdyp = (wyp - wc) * iycj, dwdym = (wc - wym) * iycjm;
      double dwdzp = (wp - wc) * izfkp, dwdzm = (wc - wm) * izfk;
      cc = visc * (dwdxp - dwdxm) * ixfi + visc * (dwdyp - dwdym) * iyfj + visc * (dwdzp - dwdzm) * izck;
    }
    // implicit diffusion: the diffusion is kept apart, dudtrk is the advection
    // alone (rk.f90:50-59)
    double ad = 0.0, bd = 0.0, cd = 0.0;
    if (IMP) { ad = a; bd = b; cd = cc; a = 0.0; b = 0.0; cc = 0.0; }
    {
      double uuip = 0.25 * (uc * dyfj * dzfk + uxp * dyfj * dzfk) * uxp;
      double uuim = 0.25 * (uc * dyfj * dzfk + uxm * dyfj * dzfk) * uxm;
      double uvjp = 0.25 * (vc * dxfi * dzfk + vxp * dxfip * dzfk) * uyp;
      double uvjm = 0.25 * (vym * dxfi * dzfk + v[ym + 1] * dxfip * dzfk) * uym;
      double uwkp = 0.25 * (wc * dxfi * dyfj + wxp * dxfip * dyfj) * up;
      double uwkm = 0.25 * (wm * dxfi * dyfj + w[zm + 1] * dxfip * dyfj) * um;
      a = a + ((-uuip + uuim) + (-uvjp + uvjm) + (-uwkp + uwkm)) * (ixci * iyfj * izfk);
      double uvip = 0.25 * (uc * dyfj * dzfk + uyp * dyfjp * dzfk) * vxp;
      double uvim = 0.25 * (uxm * dyfj * dzfk + u[xm + sx] * dyfjp * dzfk) * vxm;
      double vvjp = 0.25 * (vc * dxfi * dzfk + vyp * dxfi * dzfk) * vyp;
      double vvjm = 0.25 * (vc * dxfi * dzfk + vym * dxfi * dzfk) * vym;
      double wvkp = 0.25 * (wc * dyfj * dxfi + wyp * dyfjp * dxfi) * vp;
      double wvkm = 0.25 * (wm * dyfj * dxfi + w[zm + sx] * dyfjp * dxfi) * vm;
      b = b + ((-uvip + uvim) + (-vvjp + vvjm) + (-wvkp + wvkm)) * (ixfi * iycj * izfk);
      double uwip = 0.25 * (uc * dzfk * dyfj + up * dzfkp * dyfj) * wxp;
      double uwim = 0.25 * (uxm * dzfk * dyfj + u[xm + sz] * dzfkp * dyfj) * wxm;
      double vwjp = 0.25 * (vc * dzfk * dxfi + vp * dzfkp * dxfi) * wyp;
      double vwjm = 0.25 * (vym * dzfk * dxfi + v[ym + sz] * dzfkp * dxfi) * wym;
      double wwkp = 0.25 * (wc * dxfi * dyfj + wp * dxfi * dyfj) * wp;
      double wwkm = 0.25 * (wc * dxfi * dyfj + wm * dxfi * dyfj) * wm;
      cc = cc + ((-uwip + uwim) + (-vwjp + vwjm) + (-wwkp + wwkm)) * (ixfi * iyfj * izck);
    }
    double un_ = uc + f1 * a + f2 * dtu[q];
    double vn_ = vc + f1 * b + f2 * dtv[q];
    double wn_ = wc + f1 * cc + f2 * dtw[q];
    if (IMP) { un_ = un_ + f12 * ad; vn_ = vn_ + f12 * bd; wn_ = wn_ + f12 * cd; }  // rk.f90:75-79
    un_ = un_ + f12 * (-(p[xp] - pc) * ixci + bfx);
    vn_ = vn_ + f12 * (-(p[yp] - pc) * iycj + bfy);
    wn_ = wn_ + f12 * (-(pz - pc) * izck + bfz);
    if (BUOY) {
      const double sc = s1[c];
      un_ = un_ - f12 * gbx * 0.5 * (s1[xp] + sc);
      vn_ = vn_ - f12 * gby * 0.5 * (s1[yp] + sc);
      wn_ = wn_ - f12 * gbz * 0.5 * (s1[zp] + sc);
    }
    dtu[q] = a;
    dtv[q] = b;
    dtw[q] = cc;
    if (IMP) {  // the Helmholtz right-hand side (rk.f90:112-127)
      un_ = un_ - 0.5 * f12 * ad;
      vn_ = vn_ - 0.5 * f12 * bd;
      wn_ = wn_ - 0.5 * f12 * cd;
    }
    un[c] = un_;
    vn[c] = vn_;
    wn[c] = wn_;
    um = uc; uc = up;
    vm = vc; vc = vp;
    wm = wc; wc = wp;
    pc = pz;
  }
}

// The k-march with the planes staged in LDS: a workgroup of 64 x BY
// threads owns a 64 x BY column tile and walks KB planes; u, v, w, p of
// planes k-1, k, k+1 over the tile plus a one-cell halo sit in a ring of
// three LDS planes, so every in-plane and cross-plane neighbour is an LDS
// read and each plane is fetched from memory once per tile, with row-wise
// coalesced loads.  The loads of plane k+2 are issued before plane k is
// computed and land in the ring slot of plane k-1 after it, so their
// latency overlaps the arithmetic.
constexpr int LX = 66;  // row pitch of an LDS plane: 64 + 2 halo cells
template <bool NT, typename T>
__device__ __forceinline__ T ld_nt(const T *p) {
  if (NT) return __builtin_nontemporal_load(p);
  return *p;
}
template <bool NT, typename T>
__device__ __forceinline__ void st_nt(T *p, T v) {
  if (NT) __builtin_nontemporal_store(v, p);
  else *p = v;
}
// R: arithmetic of the momentum right-hand side (advection and diffusion)
// and storage type of the old tendencies.  double: the reference's FP64,
// in its expression order.  float (mixed precision): the right-hand side is
// evaluated in FP32 on offsets from the cell's own values -- every
// neighbour value enters as (neighbour - centre), formed in FP64 and
// rounded once -- with the advective flux differences rearranged so that
// their O(u^2) parts cancel exactly, e.g. for the x-flux of u
//   (uc+u_m) u_m - (uc+u_p) u_p = (d_m - d_p) (3 uc + d_m + d_p),
// d = u_nb - uc; rounding full-size velocities to FP32 would instead put
// an error of |u|^2 * 2^-24 into terms of size |u| |du|.  The RK3
// accumulation u + f1*rhs + f2*rhs_old, the pressure gradient and the
// body forces run in FP64.
// NTS: the read-once / write-once streams (dudtrko in and out, the new
// velocities) carry the non-temporal hint, leaving the caches to the
// stencil's planes
template <typename R, bool BUOY, bool IMP, int KB, int BY, bool NTS>
__global__ __launch_bounds__(64 * BY) void k_rk_mom_lds(Geo g, Metrics m, const double *__restrict__ u,
                                                        const double *__restrict__ v, const double *__restrict__ w,
                                                        const double *__restrict__ p, double *__restrict__ un,
                                                        double *__restrict__ vn, double *__restrict__ wn,
                                                        R *__restrict__ dtu, R *__restrict__ dtv,
                                                        R *__restrict__ dtw, double f1, double f2, double f12,
                                                        double visc_, double bfx, double bfy, double bfz,
                                                        const double *__restrict__ s1, double gbx, double gby,
                                                        double gbz) {
  constexpr bool F32 = sizeof(R) == 4;
  constexpr int NT = 64 * BY, PL = LX * (BY + 2), NL = (PL + NT - 1) / NT;
  __shared__ double sh[4][3][PL];  // field (u, v, w, p) x ring slot x plane
  int bx, by, bz;
  tile_of((g.nx + 63) / 64, (g.ny + BY - 1) / BY, bx, by, bz);
  // a wavefront is one tile row: ty (and with it j and the y metrics) is
  // wave-uniform, held in scalar registers
  const int tx = threadIdx.x, ty = __builtin_amdgcn_readfirstlane(threadIdx.y), tid = ty * 64 + tx;
  const int i0 = bx * 64 + 1, j0 = by * BY + 1;
  const int i = i0 + tx, j = j0 + ty;
  const int k0 = bz * KB + 1;
  const int k1 = min(k0 + KB - 1, g.nz);
  const long sx = g.sx, sy = g.sy, sz = sx * sy;
  // this thread's share of a plane load: LDS offset and global offset
  // within the plane (-1: outside the array)
  int lo[NL];
  long go[NL];
#pragma unroll
  for (int r = 0; r < NL; r++) {
    const int e = tid + r * NT;
    const int row = e / LX, col = e - row * LX;
    const int gi = i0 - 1 + col, gj = j0 - 1 + row;
    lo[r] = e;
    go[r] = (e < PL && gi <= g.nx + 1 && gj <= g.ny + 1) ? gidx(gi, gj, 0, sx, sy) : -1;
  }
  auto load_plane = [&](int k, double (&reg)[4][NL]) {
    const long kz = (long)k * sz;
#pragma unroll
    for (int r = 0; r < NL; r++) {
      if (go[r] >= 0) {
        reg[0][r] = u[kz + go[r]];
        reg[1][r] = v[kz + go[r]];
        reg[2][r] = w[kz + go[r]];
        reg[3][r] = p[kz + go[r]];
      }
    }
  };
  auto store_plane = [&](int slot, const double (&reg)[4][NL]) {
#pragma unroll
    for (int r = 0; r < NL; r++)
      if (go[r] >= 0)
#pragma unroll
        for (int f = 0; f < 4; f++) sh[f][slot][lo[r]] = reg[f][r];
  };
  {
    double reg[4][NL];
    for (int k = k0 - 1; k <= k0 + 1 && k <= g.nz + 1; k++) {
      load_plane(k, reg);
      store_plane(k % 3, reg);
    }
  }
  __syncthreads();
  const bool live = i <= g.nx && j <= g.ny;
  R dxfi = 0, dxfip = 0, dyfj = 0, dyfjp = 0, ixci = 0, ixcim = 0, ixfi = 0, ixfip = 0;
  R iycj = 0, iycjm = 0, iyfj = 0, iyfjp = 0;
  double ixci_d = 0, iycj_d = 0;  // the pressure gradient's, FP64
  if (live) {
    dxfi = (R)m.dxf[i]; dxfip = (R)m.dxf[i + 1]; dyfj = (R)m.dyf[j]; dyfjp = (R)m.dyf[j + 1];
    ixci_d = m.idxc[i]; iycj_d = m.idyc[j];
    ixci = (R)ixci_d; ixcim = (R)m.idxc[i - 1]; ixfi = (R)m.idxf[i]; ixfip = (R)m.idxf[i + 1];
    iycj = (R)iycj_d; iycjm = (R)m.idyc[j - 1]; iyfj = (R)m.idyf[j]; iyfjp = (R)m.idyf[j + 1];
  }
  const R visc = (R)visc_, qr = (R)0.25;
  const int L = (ty + 1) * LX + tx + 1;  // this cell in an LDS plane
  const int LXP = L + 1, LXM = L - 1, LYP = L + LX, LYM = L - LX;
  long c = gidx(i, j, k0, sx, sy);
  long q = dto_index(g, i, j, k0);
  const long qs = 3L * g.nx * g.ny;
  for (int k = k0; k <= k1; k++, c += sz, q += qs) {
    double nxt[4][NL];
    const bool pre = k + 2 <= k1 + 1;
    if (pre) load_plane(k + 2, nxt);
    const int sm_ = (k + 2) % 3, s0 = k % 3, sp_ = (k + 1) % 3;  // slots of k-1, k, k+1
    if (live) {
      const double *U0 = sh[0][s0], *V0 = sh[1][s0], *W0 = sh[2][s0], *P0 = sh[3][s0];
      const double *Um = sh[0][sm_], *Vm = sh[1][sm_], *Wm = sh[2][sm_];
      const double *Up = sh[0][sp_], *Vp = sh[1][sp_], *Wp = sh[2][sp_], *Pp = sh[3][sp_];
      const double uc_ = U0[L], vc_ = V0[L], wc_ = W0[L], pc = P0[L], pz = Pp[L];
      const double izck_d = m.idzc[k];
      const R dzfk = (R)m.dzf[k], dzfkp = (R)m.dzf[k + 1];
      const R izck = (R)izck_d, izckm = (R)m.idzc[k - 1], izfk = (R)m.idzf[k], izfkp = (R)m.idzf[k + 1];
      R a, b, cc;
      R ad = 0, bd = 0, cd = 0;
      if constexpr (F32) {
        // offsets from the centre values, one rounding each
        auto d = [](double nb, double ctr) { return (float)(nb - ctr); };
        const float uc = (float)uc_, vc = (float)vc_, wc = (float)wc_;
        const float uxp = d(U0[LXP], uc_), uxm = d(U0[LXM], uc_), uyp = d(U0[LYP], uc_), uym = d(U0[LYM], uc_);
        const float uzp = d(Up[L], uc_), uzm = d(Um[L], uc_);
        const float uxmyp = d(U0[LXM + LX], uc_), uxmzp = d(Up[LXM], uc_);
        const float vxp = d(V0[LXP], vc_), vxm = d(V0[LXM], vc_), vyp = d(V0[LYP], vc_), vym = d(V0[LYM], vc_);
        const float vzp = d(Vp[L], vc_), vzm = d(Vm[L], vc_);
        const float vymxp = d(V0[LYM + 1], vc_), vymzp = d(Vp[LYM], vc_);
        const float wxp = d(W0[LXP], wc_), wxm = d(W0[LXM], wc_), wyp = d(W0[LYP], wc_), wym = d(W0[LYM], wc_);
        const float wzp = d(Wp[L], wc_), wzm = d(Wm[L], wc_);
        const float wzmxp = d(Wm[LXP], wc_), wzmyp = d(Wm[LYP], wc_);
        // diffusion, momx_d / momy_d / momz_d (mom.f90:221-327)
        a = visc * (uxp * ixfip + uxm * ixfi) * ixci + visc * (uyp * iycj + uym * iycjm) * iyfj +
            visc * (uzp * izck + uzm * izckm) * izfk;
        b = visc * (vxp * ixci + vxm * ixcim) * ixfi + visc * (vyp * iyfjp + vym * iyfj) * iycj +
            visc * (vzp * izck + vzm * izckm) * izfk;
        cc = visc * (wxp * ixci + wxm * ixcim) * ixfi + visc * (wyp * iycj + wym * iycjm) * iyfj +
             visc * (wzp * izfkp + wzm * izfk) * izck;
        if (IMP) { ad = a; bd = b; cd = cc; a = 0; b = 0; cc = 0; }
        // volume-weighted advection, momx/y/z_a_vv (mom.f90:112-219): each
        // pair of face fluxes F_m u_m - F_p u_p as uc (F_m - F_p) + F_m d_m - F_p d_p
        {
          const float sxf = dxfi + dxfip, syf = dyfj + dyfjp, szf = dzfk + dzfkp;
          const float fuu = qr * dyfj * dzfk * (uxm - uxp) * (3.f * uc + uxm + uxp);
          const float Vp_ = vc * sxf + vxp * dxfip, Vm_ = vc * sxf + vym * dxfi + vymxp * dxfip;
          const float fuv = qr * dzfk * (uc * (vym * dxfi + (vymxp - vxp) * dxfip) + Vm_ * uym - Vp_ * uyp);
          const float Wp_ = wc * sxf + wxp * dxfip, Wm_ = wc * sxf + wzm * dxfi + wzmxp * dxfip;
          const float fuw = qr * dyfj * (uc * (wzm * dxfi + (wzmxp - wxp) * dxfip) + Wm_ * uzm - Wp_ * uzp);
          a = a + (fuu + fuv + fuw) * (ixci * iyfj * izfk);
          const float Up2 = uc * syf + uyp * dyfjp, Um2 = uc * syf + uxm * dyfj + uxmyp * dyfjp;
          const float fvu = qr * dzfk * (vc * (uxm * dyfj + (uxmyp - uyp) * dyfjp) + Um2 * vxm - Up2 * vxp);
          const float fvv = qr * dxfi * dzfk * (vym - vyp) * (3.f * vc + vym + vyp);
          const float Wp2 = wc * syf + wyp * dyfjp, Wm2 = wc * syf + wzm * dyfj + wzmyp * dyfjp;
          const float fvw = qr * dxfi * (vc * (wzm * dyfj + (wzmyp - wyp) * dyfjp) + Wm2 * vzm - Wp2 * vzp);
          b = b + (fvu + fvv + fvw) * (ixfi * iycj * izfk);
          const float Up3 = uc * szf + uzp * dzfkp, Um3 = uc * szf + uxm * dzfk + uxmzp * dzfkp;
          const float fwu = qr * dyfj * (wc * (uxm * dzfk + (uxmzp - uzp) * dzfkp) + Um3 * wxm - Up3 * wxp);
          const float Vp3 = vc * szf + vzp * dzfkp, Vm3 = vc * szf + vym * dzfk + vymzp * dzfkp;
          const float fwv = qr * dxfi * (wc * (vym * dzfk + (vymzp - vzp) * dzfkp) + Vm3 * wym - Vp3 * wyp);
          const float fww = qr * dxfi * dyfj * (wzm - wzp) * (3.f * wc + wzm + wzp);
          cc = cc + (fwu + fwv + fww) * (ixfi * iyfj * izck);
        }
      } else {
        const double uc = uc_, vc = vc_, wc = wc_;
        const double um = Um[L], vm = Vm[L], wm = Wm[L];
        const double up = Up[L], vp = Vp[L], wp = Wp[L];
        const double uxp = U0[LXP], uxm = U0[LXM], uyp = U0[LYP], uym = U0[LYM];
        const double vxp = V0[LXP], vxm = V0[LXM], vyp = V0[LYP], vym = V0[LYM];
        const double wxp = W0[LXP], wxm = W0[LXM], wyp = W0[LYP], wym = W0[LYM];
        {
          double dudxp = (uxp - uc) * ixfip, dudxm = (uc - uxm) * ixfi;
          double dudyp = (uyp - uc) * iycj, dudym = (uc - uym) * iycjm;
          double dudzp = (up - uc) * izck, dudzm = (uc - um) * izckm;
          a = visc * (dudxp - dudxm) * ixci + visc * (dudyp - dudym) * iyfj + visc * (dudzp - dudzm) * izfk;
          double dvdxp = (vxp - vc) * ixci, dvdxm = (vc - vxm) * ixcim;
          double dvdyp = (vyp - vc) * iyfjp, dvdym = (vc - vym) * iyfj;
          double dvdzp = (vp - vc) * izck, dvdzm = (vc - vm) * izckm;
          b = visc * (dvdxp - dvdxm) * ixfi + visc * (dvdyp - dvdym) * iycj + visc * (dvdzp - dvdzm) * izfk;
          double dwdxp = (wxp - wc) * ixci, dwdxm = (wc - wxm) * ixcim;
          double dwdyp = (wyp - wc) * iycj, dwdym = (wc - wym) * iycjm;
          double dwdzp = (wp - wc) * izfkp, dwdzm = (wc - wm) * izfk;
          cc = visc * (dwdxp - dwdxm) * ixfi + visc * (dwdyp - dwdym) * iyfj + visc * (dwdzp - dwdzm) * izck;
        }
        // implicit diffusion: the diffusion is kept apart, dudtrk is the advection
        // alone (rk.f90:50-59)
        if (IMP) { ad = a; bd = b; cd = cc; a = 0; b = 0; cc = 0; }
        {
          double uuip = qr * (uc * dyfj * dzfk + uxp * dyfj * dzfk) * uxp;
          double uuim = qr * (uc * dyfj * dzfk + uxm * dyfj * dzfk) * uxm;
          double uvjp = qr * (vc * dxfi * dzfk + vxp * dxfip * dzfk) * uyp;
          double uvjm = qr * (vym * dxfi * dzfk + V0[LYM + 1] * dxfip * dzfk) * uym;
          double uwkp = qr * (wc * dxfi * dyfj + wxp * dxfip * dyfj) * up;
          double uwkm = qr * (wm * dxfi * dyfj + Wm[LXP] * dxfip * dyfj) * um;
          a = a + ((-uuip + uuim) + (-uvjp + uvjm) + (-uwkp + uwkm)) * (ixci * iyfj * izfk);
          double uvip = qr * (uc * dyfj * dzfk + uyp * dyfjp * dzfk) * vxp;
          double uvim = qr * (uxm * dyfj * dzfk + U0[LXM + LX] * dyfjp * dzfk) * vxm;
          double vvjp = qr * (vc * dxfi * dzfk + vyp * dxfi * dzfk) * vyp;
          double vvjm = qr * (vc * dxfi * dzfk + vym * dxfi * dzfk) * vym;
          double wvkp = qr * (wc * dyfj * dxfi + wyp * dyfjp * dxfi) * vp;
          double wvkm = qr * (wm * dyfj * dxfi + Wm[LYP] * dyfjp * dxfi) * vm;
          b = b + ((-uvip + uvim) + (-vvjp + vvjm) + (-wvkp + wvkm)) * (ixfi * iycj * izfk);
          double uwip = qr * (uc * dzfk * dyfj + up * dzfkp * dyfj) * wxp;
          double uwim = qr * (uxm * dzfk * dyfj + Up[LXM] * dzfkp * dyfj) * wxm;
          double vwjp = qr * (vc * dzfk * dxfi + vp * dzfkp * dxfi) * wyp;
          double vwjm = qr * (vym * dzfk * dxfi + Vp[LYM] * dzfkp * dxfi) * wym;
          double wwkp = qr * (wc * dxfi * dyfj + wp * dxfi * dyfj) * wp;
          double wwkm = qr * (wc * dxfi * dyfj + wm * dxfi * dyfj) * wm;
          cc = cc + ((-uwip + uwim) + (-vwjp + vwjm) + (-wwkp + wwkm)) * (ixfi * iyfj * izck);
        }
      }
      double un_ = uc_ + f1 * (double)a + f2 * (double)ld_nt<NTS>(dtu + q);
      double vn_ = vc_ + f1 * (double)b + f2 * (double)ld_nt<NTS>(dtv + q);
      double wn_ = wc_ + f1 * (double)cc + f2 * (double)ld_nt<NTS>(dtw + q);
      if (IMP) {  // rk.f90:75-79
        un_ = un_ + f12 * (double)ad; vn_ = vn_ + f12 * (double)bd; wn_ = wn_ + f12 * (double)cd;
      }
      un_ = un_ + f12 * (-(P0[LXP] - pc) * ixci_d + bfx);
      vn_ = vn_ + f12 * (-(P0[LYP] - pc) * iycj_d + bfy);
      wn_ = wn_ + f12 * (-(pz - pc) * izck_d + bfz);
      if (BUOY) {
        const double sc = s1[c];
        un_ = un_ - f12 * gbx * 0.5 * (s1[c + 1] + sc);
        vn_ = vn_ - f12 * gby * 0.5 * (s1[c + sx] + sc);
        wn_ = wn_ - f12 * gbz * 0.5 * (s1[c + sz] + sc);
      }
      st_nt<NTS>(dtu + q, a);
      st_nt<NTS>(dtv + q, b);
      st_nt<NTS>(dtw + q, cc);
      if (IMP) {  // the Helmholtz right-hand side (rk.f90:112-127)
        un_ = un_ - 0.5 * f12 * (double)ad;
        vn_ = vn_ - 0.5 * f12 * (double)bd;
        wn_ = wn_ - 0.5 * f12 * (double)cd;
      }
      st_nt<NTS>(un + c, un_);
      st_nt<NTS>(vn + c, vn_);
      st_nt<NTS>(wn + c, wn_);
    }
    __syncthreads();  // plane k-1's slot is free
    if (pre) store_plane(sm_, nxt);
    __syncthreads();
  }
}

// rk_scal (rk.f90:130-204) with scal_d / scal_a (scal.f90:89-155) fused:
// reads s (ghosted) and u, v, w; writes the updated interior into sn (the
// other buffer of the pair) and dsdtrk into dso.
// IMP (implicit diffusion, rk.f90:151-202): 0 off; 1 the diffusion kept out
// of dsdtrk, half of it left in the update; 2 the same with the half taken
// off later (a forced scalar's bulk mean is taken before it): the diffusion
// goes to ddo
template <int IMP>
__global__ __launch_bounds__(TX *TY) void k_rk_scal(Geo g, Metrics m, const double *__restrict__ u,
                                                     const double *__restrict__ v, const double *__restrict__ w,
                                                     const double *__restrict__ s, double *__restrict__ sn,
                                                     double *__restrict__ dso, double alpha, double f1, double f2,
                                                     double f12, double src, double *__restrict__ ddo) {
  int bx, by, bz;
  tile_of((g.nx + TX - 1) / TX, (g.ny + TY - 1) / TY, bx, by, bz);
  const int i = bx * TX + threadIdx.x + 1;
  const int j = by * TY + threadIdx.y + 1;
  const int k = bz + 1;
  if (i > g.nx || j > g.ny) return;
  const long sx = g.sx, sy = g.sy;
  const long c = gidx(i, j, k, sx, sy);
  const long xp = c + 1, xm = c - 1, yp = c + sx, ym = c - sx, zp = c + sx * sy, zm = c - sx * sy;
  const double sc = s[c], sxp = s[xp], sxm = s[xm], syp = s[yp], sym = s[ym], szp = s[zp], szm = s[zm];
  const double ixfi = m.idxf[i], iyfj = m.idyf[j], izfk = m.idzf[k];
  double d = alpha * ((sxp - sc) * m.idxc[i] - (sc - sxm) * m.idxc[i - 1]) * ixfi +
             alpha * ((syp - sc) * m.idyc[j] - (sc - sym) * m.idyc[j - 1]) * iyfj +
             alpha * ((szp - sc) * m.idzc[k] - (sc - szm) * m.idzc[k - 1]) * izfk;
  const double usim = 0.5 * (sxm + sc) * u[xm], usip = 0.5 * (sxp + sc) * u[c];
  const double vsjm = 0.5 * (sym + sc) * v[ym], vsjp = 0.5 * (syp + sc) * v[c];
  const double wskm = 0.5 * (szm + sc) * w[zm], wskp = 0.5 * (szp + sc) * w[c];
  double dd = 0.0;
  if (IMP) { dd = d; d = 0.0; }
  d = d + (-usip + usim) * ixfi + (-vsjp + vsjm) * iyfj + (-wskp + wskm) * izfk;
  const long q = (long)(i - 1) + (long)g.nx * ((long)(j - 1) + (long)g.ny * (long)(k - 1));
  double snv = sc + f1 * d + f2 * dso[q] + f12 * src;
  if (IMP) snv = snv + f12 * dd;
  if (IMP == 1) snv = snv - 0.5 * f12 * dd;
  if (IMP == 2) ddo[q] = dd;
  sn[c] = snv;
  dso[q] = d;
}

__global__ __launch_bounds__(TX *TY) void k_fillps(Geo g, Metrics m, int fa, const double *__restrict__ u,
                                                    const double *__restrict__ v, const double *__restrict__ w,
                                                    double *__restrict__ pp, double idtrk, const double *rx0,
                                                    const double *rx1, const double *ry0, const double *ry1,
                                                    const double *rz0, const double *rz1) {
  int bx, by, bz;
  tile_of((g.nx + TX - 1) / TX, (g.ny + TY - 1) / TY, bx, by, bz);
  const int i = bx * TX + threadIdx.x + 1;
  const int j = by * TY + threadIdx.y + 1;
  const int k = bz + 1;
  if (i > g.nx || j > g.ny) return;
  const long sx = g.sx, sy = g.sy;
  const long c = gidx(i, j, k, sx, sy);
  const double dw = w[c] - w[c - sx * sy], dv = v[c] - v[c - sx], du = u[c] - u[c - 1];
  const double dxf = m.dxf[i], dyf = m.dyf[j], dzf = m.dzf[k];
  double s;
  switch (fa) {
  case 3: s = dw * m.idzf[k] * dxf * dyf + dv * dxf + du * dyf; break;
  case 2: s = dw * dxf + dv * m.idyf[j] * dxf * dzf + du * dzf; break;
  case 1: s = dw * dyf + dv * dzf + du * m.idxf[i] * dyf * dzf; break;
  default: s = dw * dxf * dyf + dv * dxf * dzf + du * dyf * dzf; break;
  }
  double r = s * idtrk;
  // updt_rhs, in the reference's order x0,x1,y0,y1,z0,z1
  if (rx0 && i == 1) r += rx0[(j - 1) + (long)g.ny * (k - 1)];
  if (rx1 && i == g.nx) r += rx1[(j - 1) + (long)g.ny * (k - 1)];
  if (ry0 && j == 1) r += ry0[(i - 1) + (long)g.nx * (k - 1)];
  if (ry1 && j == g.ny) r += ry1[(i - 1) + (long)g.nx * (k - 1)];
  if (rz0 && k == 1) r += rz0[(i - 1) + (long)g.nx * (j - 1)];
  if (rz1 && k == g.nz) r += rz1[(i - 1) + (long)g.nx * (j - 1)];
  st_nt<true>(pp + c, r);  // written once, read by the next kernel
}

// correc over the whole ghosted box; cells outside the reference's loop
// ranges (i <= nx for u, j <= ny for v, k <= nz for w) are copied
__global__ __launch_bounds__(TX *TY) void k_correc(Geo g, Metrics m, const double *__restrict__ pp,
                                                    const double *__restrict__ u, const double *__restrict__ v,
                                                    const double *__restrict__ w, double *__restrict__ un,
                                                    double *__restrict__ vn, double *__restrict__ wn, double dtrk,
                                                    double *__restrict__ p) {
  int bx, by, bz;
  tile_of((g.nx + 2 + TX - 1) / TX, (g.ny + 2 + TY - 1) / TY, bx, by, bz);
  const int i = bx * TX + threadIdx.x;
  const int j = by * TY + threadIdx.y;
  const int k = bz;
  if (i > g.nx + 1 || j > g.ny + 1) return;
  const long sx = g.sx, sy = g.sy;
  const long c = gidx(i, j, k, sx, sy);
  const double pc = pp[c];
  double uu = u[c], vv = v[c], ww = w[c];
  if (i <= g.nx) uu = uu - dtrk * (pp[c + 1] - pc) * m.idxc[i];
  if (j <= g.ny) vv = vv - dtrk * (pp[c + sx] - pc) * m.idyc[j];
  if (k <= g.nz) ww = ww - dtrk * (pp[c + sx * sy] - pc) * m.idzc[k];
  un[c] = uu;
  vn[c] = vv;
  wn[c] = ww;
  if (p && i >= 1 && i <= g.nx && j >= 1 && j <= g.ny && k >= 1 && k <= g.nz) p[c] = p[c] + pc;
}

__global__ __launch_bounds__(TX *TY) void k_updt_pressure(Geo g, const double *__restrict__ pp,
                                                           double *__restrict__ p) {
  const int i = blockIdx.x * TX + threadIdx.x + 1;
  const int j = blockIdx.y * TY + threadIdx.y + 1;
  const int k = blockIdx.z + 1;
  if (i > g.nx || j > g.ny) return;
  const long c = gidx(i, j, k, g.sx, g.sy);
  p[c] = p[c] + pp[c];
}

// ---- reductions ----
__device__ __forceinline__ double wave_sum(double x) {
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
  return x;
}
__device__ __forceinline__ double wave_max(double x) {
  for (int o = 32; o > 0; o >>= 1) x = fmax(x, __shfl_xor(x, o, 64));
  return x;
}

__global__ __launch_bounds__(TX *TY) void k_chkdiv(Geo g, Metrics m, const double *__restrict__ u,
                                                    const double *__restrict__ v, const double *__restrict__ w,
                                                    double vol, double *psum, double *pmax) {
  __shared__ double ss[TX * TY / 64], sm[TX * TY / 64];
  const int i = blockIdx.x * TX + threadIdx.x + 1;
  const int j = blockIdx.y * TY + threadIdx.y + 1;
  const int k = blockIdx.z + 1;
  double s = 0.0, mx = 0.0;
  if (i <= g.nx && j <= g.ny) {
    const long sx = g.sx, sy = g.sy;
    const long c = gidx(i, j, k, sx, sy);
    double div = (w[c] - w[c - sx * sy]) / m.dzf[k] + (v[c] - v[c - sx]) / m.dyf[j] + (u[c] - u[c - 1]) / m.dxf[i];
    mx = fabs(div);
    s = div * m.dxf[i] * m.dyf[j] * m.dzf[k] / vol;
  }
  s = wave_sum(s);
  mx = wave_max(mx);
  const int t = threadIdx.x + TX * threadIdx.y, wv = t / 64;
  if ((t & 63) == 0) { ss[wv] = s; sm[wv] = mx; }
  __syncthreads();
  if (t == 0) {
    double a = 0.0, b = 0.0;
    for (int q = 0; q < TX * TY / 64; q++) { a += ss[q]; b = fmax(b, sm[q]); }
    const long part = blockIdx.x + (long)gridDim.x * (blockIdx.y + (long)gridDim.y * blockIdx.z);
    psum[part] = a;
    pmax[part] = b;
  }
}

__global__ __launch_bounds__(TX *TY) void k_chkdt(Geo g, Metrics m, const double *__restrict__ u,
                                                   const double *__restrict__ v, const double *__restrict__ w,
                                                   double *pmax) {
  __shared__ double sm[TX * TY / 64];
  const int i = blockIdx.x * TX + threadIdx.x + 1;
  const int j = blockIdx.y * TY + threadIdx.y + 1;
  const int k = blockIdx.z + 1;
  double dti = 0.0;
  if (i <= g.nx && j <= g.ny) {
    const long sx = g.sx, sy = g.sy, sz = sx * sy;
    const long c = gidx(i, j, k, sx, sy);
    double ux = fabs(u[c]);
    double vx = 0.25 * fabs(v[c] + v[c - sx] + v[c + 1] + v[c + 1 - sx]);
    double wx = 0.25 * fabs(w[c] + w[c - sz] + w[c + 1] + w[c + 1 - sz]);
    double dtix = ux / m.dxc[i] + vx / m.dyf[j] + wx / m.dzf[k];
    double uy = 0.25 * fabs(u[c] + u[c + sx] + u[c - 1 + sx] + u[c - 1]);
    double vy = fabs(v[c]);
    double wy = 0.25 * fabs(w[c] + w[c + sx] + w[c + sx - sz] + w[c - sz]);
    double dtiy = uy / m.dxf[i] + vy / m.dyc[j] + wy / m.dzf[k];
    double uz = 0.25 * fabs(u[c] + u[c - 1] + u[c - 1 + sz] + u[c + sz]);
    double vz = 0.25 * fabs(v[c] + v[c - sx] + v[c - sx + sz] + v[c + sz]);
    double wz = fabs(w[c]);
    double dtiz = uz / m.dxf[i] + vz / m.dyf[j] + wz / m.dzc[k];
    dti = fmax(dtix, fmax(dtiy, dtiz));
  }
  dti = wave_max(dti);
  const int t = threadIdx.x + TX * threadIdx.y, wv = t / 64;
  if ((t & 63) == 0) sm[wv] = dti;
  __syncthreads();
  if (t == 0) {
    double b = 0.0;
    for (int q = 0; q < TX * TY / 64; q++) b = fmax(b, sm[q]);
    pmax[blockIdx.x + (long)gridDim.x * (blockIdx.y + (long)gridDim.y * blockIdx.z)] = b;
  }
}

// partial sums of s*dxf*dyf*dzf (bulk_mean_s scal.f90:157-180)
__global__ __launch_bounds__(TX *TY) void k_volsum(Geo g, Metrics m, const double *__restrict__ s, double *psum) {
  __shared__ double ss[TX * TY / 64];
  const int i = blockIdx.x * TX + threadIdx.x + 1;
  const int j = blockIdx.y * TY + threadIdx.y + 1;
  const int k = blockIdx.z + 1;
  double a = 0.0;
  if (i <= g.nx && j <= g.ny) a = s[gidx(i, j, k, g.sx, g.sy)] * m.dxf[i] * m.dyf[j] * m.dzf[k];
  a = wave_sum(a);
  const int t = threadIdx.x + TX * threadIdx.y;
  if ((t & 63) == 0) ss[t / 64] = a;
  __syncthreads();
  if (t == 0) {
    double b = 0.0;
    for (int q = 0; q < TX * TY / 64; q++) b += ss[q];
    psum[blockIdx.x + (long)gridDim.x * (blockIdx.y + (long)gridDim.y * blockIdx.z)] = b;
  }
}

// bulk_forcing_s scal.f90:182-199 with f = scalf - (volume sum)/vol_all read
// from the device; the first thread also records f
// s += f; with dd (implicit diffusion of a forced scalar) s - h*dd first
__global__ __launch_bounds__(TX *TY) void k_bulk_force(Geo g, double *__restrict__ s, const double *__restrict__ red,
                                                        double scalf, double ivol, double *fout,
                                                        const double *__restrict__ dd, double h) {
  const int i = blockIdx.x * TX + threadIdx.x + 1;
  const int j = blockIdx.y * TY + threadIdx.y + 1;
  const int k = blockIdx.z + 1;
  const double f = scalf - red[0] * ivol;
  if (fout && i == 1 && j == 1 && k == 1) fout[0] = f;
  if (i <= g.nx && j <= g.ny) {
    const long c = gidx(i, j, k, g.sx, g.sy);
    double v = s[c];
    if (dd) v = v - h * dd[(long)(i - 1) + (long)g.nx * ((long)(j - 1) + (long)g.ny * (long)(k - 1))];
    s[c] = v + f;
  }
}

__global__ __launch_bounds__(256) void k_reduce_final(const double *psum, const double *pmax, int n, double *out) {
  __shared__ double ss[4], sm[4];
  double s = 0.0, mx = 0.0;
  for (int q = threadIdx.x; q < n; q += 256) {
    if (psum) s += psum[q];
    if (pmax) mx = fmax(mx, pmax[q]);
  }
  s = wave_sum(s);
  mx = wave_max(mx);
  if ((threadIdx.x & 63) == 0) { ss[threadIdx.x / 64] = s; sm[threadIdx.x / 64] = mx; }
  __syncthreads();
  if (threadIdx.x == 0) {
    out[0] = ss[0] + ss[1] + ss[2] + ss[3];
    out[1] = fmax(fmax(sm[0], sm[1]), fmax(sm[2], sm[3]));
  }
}

dim3 interior_grid(const Geo &g) { return dim3((g.nx + TX - 1) / TX, (g.ny + TY - 1) / TY, g.nz); }
dim3 flat(dim3 g) { return dim3(g.x * g.y * g.z); }
dim3 interior_tiles(const Geo &g) { return flat(interior_grid(g)); }
dim3 ghosted_tiles(const Geo &g) { return dim3(((g.nx + 2 + TX - 1) / TX) * ((g.ny + 2 + TY - 1) / TY) * (g.nz + 2)); }

}  // namespace

namespace {
// the LDS-staged k-march forms (variants 7..16) with FP64 or FP32
// right-hand-side arithmetic
template <typename R>
void launch_rk_mom_lds(hipStream_t s, const Geo &g, const Metrics &m, const double *u, const double *v,
                       const double *w, const double *p, double *un, double *vn, double *wn, R *dtu, R *dtv,
                       R *dtw, double f1, double f2, double f12, double visc, double bfx, double bfy, double bfz,
                       const double *s1, const double gb[3], int variant, bool impdiff) {
  // 7..10: LDS-staged k-march; 7/8: 64x4 tiles, KB 16/32; 9/10: 64x8 tiles, KB 16/32;
  // 11..14: the same with non-temporal hints on the streams;
  // (capping the registers at 128 for a fourth wave per SIMD spills and measured slower)
  // 15, 16: 64x8 tiles over 64 planes, without / with the hints
  // (64x12 tiles, twelve waves per workgroup, measured 6-25 % slower)
  const bool nt = variant >= 11 && variant != 15;
  const int vi = variant >= 15 ? 4 : (variant - 7) % 4;
  const int kb = vi == 4 ? 64 : (vi % 2 ? 32 : 16), by = vi >= 2 ? 8 : 4;
  dim3 grid((unsigned)(((g.nx + 63) / 64) * ((g.ny + by - 1) / by) * ((g.nz + kb - 1) / kb)));
#define SNAC_KL(B, KB, Y)                                                                                           \
  do {                                                                                                              \
    if (nt)                                                                                                         \
      hipLaunchKernelGGL((k_rk_mom_lds<R, B, IM, KB, Y, true>), grid, dim3(64, Y), 0, s, g, m, u, v, w, p, un, vn,   \
                         wn, dtu, dtv, dtw, f1, f2, f12, visc, bfx, bfy, bfz, s1, gb[0], gb[1], gb[2]);              \
    else                                                                                                            \
      hipLaunchKernelGGL((k_rk_mom_lds<R, B, IM, KB, Y, false>), grid, dim3(64, Y), 0, s, g, m, u, v, w, p, un, vn,  \
                         wn, dtu, dtv, dtw, f1, f2, f12, visc, bfx, bfy, bfz, s1, gb[0], gb[1], gb[2]);              \
  } while (0)
#define SNAC_KL_ALL(B)                                                                                              \
  do {                                                                                                              \
    if (vi == 0) SNAC_KL(B, 16, 4);                                                                                 \
    else if (vi == 1) SNAC_KL(B, 32, 4);                                                                            \
    else if (vi == 2) SNAC_KL(B, 16, 8);                                                                            \
    else if (vi == 3) SNAC_KL(B, 32, 8);                                                                            \
    else SNAC_KL(B, 64, 8);                                                                                         \
  } while (0)
  if (impdiff) {
    constexpr bool IM = true;
    if (s1) SNAC_KL_ALL(true);
    else SNAC_KL_ALL(false);
  } else {
    constexpr bool IM = false;
    if (s1) SNAC_KL_ALL(true);
    else SNAC_KL_ALL(false);
  }
#undef SNAC_KL_ALL
#undef SNAC_KL
  SNAC_HIP(hipGetLastError());
}
}  // namespace

void launch_rk_mom_f32(hipStream_t s, const Geo &g, const Metrics &m, const double *u, const double *v,
                       const double *w, const double *p, double *un, double *vn, double *wn, float *dtu, float *dtv,
                       float *dtw, double f1, double f2, double f12, double visc, double bfx, double bfy, double bfz,
                       const double *s1, const double gb[3], int variant, bool impdiff) {
  if (variant < 7 || variant > 16) throw std::runtime_error("FP32 rk_mom: LDS variants 7..16 only");
  launch_rk_mom_lds<float>(s, g, m, u, v, w, p, un, vn, wn, dtu, dtv, dtw, f1, f2, f12, visc, bfx, bfy, bfz, s1, gb,
                           variant, impdiff);
}

void launch_rk_mom(hipStream_t s, const Geo &g, const Metrics &m, const double *u, const double *v, const double *w,
                   const double *p, double *un, double *vn, double *wn, double *dtu, double *dtv, double *dtw,
                   double f1, double f2, double f12, double visc, double bfx, double bfy, double bfz,
                   const double *s1, const double gb[3], int variant, bool impdiff) {
  if (variant >= 7) {
    launch_rk_mom_lds<double>(s, g, m, u, v, w, p, un, vn, wn, dtu, dtv, dtw, f1, f2, f12, visc, bfx, bfy, bfz, s1,
                              gb, variant, impdiff);
    return;
  }
  if (variant > 0) {
    // 1..6: k-march; (variant-1) % 3 -> KB 8/16/32, / 3 -> tile 64x4, 32x8
    const int vi = variant - 1, kb = 8 << (vi % 3), shape = vi / 3;
    const int bxs = shape == 1 ? 32 : 64, bys = 256 / bxs;
    dim3 grid((unsigned)(((g.nx + bxs - 1) / bxs) * ((g.ny + bys - 1) / bys) * ((g.nz + kb - 1) / kb)));
#define SNAC_KZ(B, KB, X, Y)                                                                                        \
  hipLaunchKernelGGL((k_rk_mom_kz<B, IM, KB, X, Y>), grid, dim3(X, Y), 0, s, g, m, u, v, w, p, un, vn, wn, dtu, dtv, dtw,  \
                     f1, f2, f12, visc, bfx, bfy, bfz, s1, gb[0], gb[1], gb[2])
#define SNAC_KZ_KB(B, X, Y)                                                                                         \
  do {                                                                                                              \
    if (kb == 8) SNAC_KZ(B, 8, X, Y);                                                                               \
    else if (kb == 16) SNAC_KZ(B, 16, X, Y);                                                                        \
    else SNAC_KZ(B, 32, X, Y);                                                                                      \
  } while (0)
#define SNAC_KZ_SHAPE(B)                                                                                            \
  do {                                                                                                              \
    if (shape == 0) SNAC_KZ_KB(B, 64, 4);                                                                           \
    else SNAC_KZ_KB(B, 32, 8);                                                                                      \
  } while (0)
    if (impdiff) {
      constexpr bool IM = true;
      if (s1) SNAC_KZ_SHAPE(true);
      else SNAC_KZ_SHAPE(false);
    } else {
      constexpr bool IM = false;
      if (s1) SNAC_KZ_SHAPE(true);
      else SNAC_KZ_SHAPE(false);
    }
#undef SNAC_KZ_SHAPE
#undef SNAC_KZ_KB
#undef SNAC_KZ
    SNAC_HIP(hipGetLastError());
    return;
  }
#define SNAC_K0(B, I)                                                                                               \
  hipLaunchKernelGGL((k_rk_mom<B, I>), interior_tiles(g), dim3(TX, TY), 0, s, g, m, u, v, w, p, un, vn, wn, dtu, dtv,  \
                     dtw, f1, f2, f12, visc, bfx, bfy, bfz, s1, gb[0], gb[1], gb[2])
  if (s1) {
    if (impdiff) SNAC_K0(true, true);
    else SNAC_K0(true, false);
  } else {
    if (impdiff) SNAC_K0(false, true);
    else SNAC_K0(false, false);
  }
#undef SNAC_K0
  SNAC_HIP(hipGetLastError());
}

void launch_rk_scal(hipStream_t s, const Geo &g, const Metrics &m, const double *u, const double *v, const double *w,
                    const double *sc, double *sn, double *dso, double alpha, double f1, double f2, double f12,
                    double src, int imp, double *ddo) {
  if (imp == 2)
    hipLaunchKernelGGL(k_rk_scal<2>, interior_tiles(g), dim3(TX, TY), 0, s, g, m, u, v, w, sc, sn, dso, alpha, f1, f2,
                       f12, src, ddo);
  else if (imp == 1)
    hipLaunchKernelGGL(k_rk_scal<1>, interior_tiles(g), dim3(TX, TY), 0, s, g, m, u, v, w, sc, sn, dso, alpha, f1, f2,
                       f12, src, ddo);
  else
    hipLaunchKernelGGL(k_rk_scal<0>, interior_tiles(g), dim3(TX, TY), 0, s, g, m, u, v, w, sc, sn, dso, alpha, f1, f2,
                       f12, src, ddo);
  SNAC_HIP(hipGetLastError());
}

void launch_volsum_part(hipStream_t s, const Geo &g, const Metrics &m, const double *f, double *psum) {
  hipLaunchKernelGGL(k_volsum, interior_grid(g), dim3(TX, TY), 0, s, g, m, f, psum);
  SNAC_HIP(hipGetLastError());
}

void launch_bulk_force(hipStream_t s, const Geo &g, double *f, const double *red, double scalf, double ivol,
                       double *fout, const double *dd, double h) {
  hipLaunchKernelGGL(k_bulk_force, interior_grid(g), dim3(TX, TY), 0, s, g, f, red, scalf, ivol, fout, dd, h);
  SNAC_HIP(hipGetLastError());
}

void launch_fillps(hipStream_t s, const Geo &g, const Metrics &m, int fa, const double *u, const double *v,
                   const double *w, double *pp, double dtrk, const double *const r[6]) {
  hipLaunchKernelGGL(k_fillps, interior_tiles(g), dim3(TX, TY), 0, s, g, m, fa, u, v, w, pp, 1.0 / dtrk, r[0], r[1], r[2],
                     r[3], r[4], r[5]);
  SNAC_HIP(hipGetLastError());
}

void launch_correc(hipStream_t s, const Geo &g, const Metrics &m, const double *pp, const double *u, const double *v,
                   const double *w, double *un, double *vn, double *wn, double dtrk, double *p) {
  hipLaunchKernelGGL(k_correc, ghosted_tiles(g), dim3(TX, TY), 0, s, g, m, pp, u, v, w, un, vn, wn, dtrk, p);
  SNAC_HIP(hipGetLastError());
}

void launch_updt_pressure(hipStream_t s, const Geo &g, const double *pp, double *p) {
  hipLaunchKernelGGL(k_updt_pressure, interior_grid(g), dim3(TX, TY), 0, s, g, pp, p);
  SNAC_HIP(hipGetLastError());
}

int reduce_nparts(const Geo &g) {
  dim3 gr = interior_grid(g);
  return (int)(gr.x * gr.y * gr.z);
}

void launch_chkdiv_part(hipStream_t s, const Geo &g, const Metrics &m, const double *u, const double *v,
                        const double *w, double vol, double *psum, double *pmax) {
  hipLaunchKernelGGL(k_chkdiv, interior_grid(g), dim3(TX, TY), 0, s, g, m, u, v, w, vol, psum, pmax);
  SNAC_HIP(hipGetLastError());
}

void launch_chkdt_part(hipStream_t s, const Geo &g, const Metrics &m, const double *u, const double *v,
                       const double *w, double *pmax) {
  hipLaunchKernelGGL(k_chkdt, interior_grid(g), dim3(TX, TY), 0, s, g, m, u, v, w, pmax);
  SNAC_HIP(hipGetLastError());
}

void launch_reduce_final(hipStream_t s, const double *psum, const double *pmax, int n, double *out) {
  hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(256), 0, s, psum, pmax, n, out);
  SNAC_HIP(hipGetLastError());
}

}  // namespace snac

namespace snac {
namespace {
// plain copy, one 16-byte element per lane over an exact grid (the fastest
// copy form measured on MI355X): the HBM reference kernel for the roofline
// and for scaling the FETCH_SIZE / WRITE_SIZE counters
__global__ __launch_bounds__(256) void k_cvt_d2f(const double *__restrict__ a, float *__restrict__ b, long n) {
  for (long e = blockIdx.x * 256L + threadIdx.x; e < n; e += (long)gridDim.x * 256) b[e] = (float)a[e];
}
__global__ __launch_bounds__(256) void k_cvt_f2d(const float *__restrict__ a, double *__restrict__ b, long n) {
  for (long e = blockIdx.x * 256L + threadIdx.x; e < n; e += (long)gridDim.x * 256) b[e] = (double)a[e];
}
__global__ __launch_bounds__(256) void k_stream_copy(const double2 *__restrict__ a, double2 *__restrict__ b, long n2) {
  const long i = blockIdx.x * 256L + threadIdx.x;
  if (i < n2) b[i] = a[i];
}
}  // namespace

void launch_cvt_d2f(hipStream_t s, const double *a, float *b, long n) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_cvt_d2f, dim3((unsigned)std::min<long>((n + 255) / 256, 8192)), dim3(256), 0, s, a, b, n);
  SNAC_HIP(hipGetLastError());
}
void launch_cvt_f2d(hipStream_t s, const float *a, double *b, long n) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_cvt_f2d, dim3((unsigned)std::min<long>((n + 255) / 256, 8192)), dim3(256), 0, s, a, b, n);
  SNAC_HIP(hipGetLastError());
}
void launch_stream_copy(hipStream_t s, const double *a, double *b, long n) {
  long n2 = n / 2;
  hipLaunchKernelGGL(k_stream_copy, dim3((unsigned)((n2 + 255) / 256)), dim3(256), 0, s, (const double2 *)a,
                     (double2 *)b, n2);
  SNAC_HIP(hipGetLastError());
}
}  // namespace snac
