// Implicit diffusion (the reference's _IMPDIFF build): the right-hand side
// of one field's Helmholtz problem (solve_impdiff_field,
// solve_helmholtz.f90:62-81, with updt_rhs bound.f90:543-585) and the
// Laplacian term of updt_pressure (updt_pressure.f90:30-45).  Both are
// HBM-bound streams over the block: one read and one write per cell.
#include "common.hpp"

namespace snac {
namespace {

constexpr int TX = 64, TY = 4;

// out = phi * alphai * w0[i] * w1[j] * w2[k] over the unknowns, plus the BC
// planes of the bounded faces in updt_rhs's order (x lo, x hi, y lo, ...);
// 0 at the cells past the unknowns (the bounded upper face of a staggered
// component), which the solve keeps decoupled
__global__ __launch_bounds__(TX *TY) void k_imp_rhs(Geo g, ImpRhs r, const double *__restrict__ phi,
                                                    double *__restrict__ out, double alphai) {
  const int i = blockIdx.x * TX + threadIdx.x + 1;
  const int j = blockIdx.y * TY + threadIdx.y + 1;
  const int k = blockIdx.z + 1;
  if (i > g.nx || j > g.ny) return;
  const long c = i + g.sx * (j + g.sy * (long)k);
  if (i > r.ext[0] || j > r.ext[1] || k > r.ext[2]) {
    out[c] = 0.0;
    return;
  }
  double v = phi[c] * alphai;
  v = v * r.w[0][i - 1];
  v = v * r.w[1][j - 1];
  v = v * r.w[2][k - 1];
  const int e0 = r.ext[0], e1 = r.ext[1];
  if (i == 1 && r.plane[0][0]) v = v + r.plane[0][0][(j - 1) + (long)e1 * (k - 1)];
  if (i == e0 && r.plane[0][1]) v = v + r.plane[0][1][(j - 1) + (long)e1 * (k - 1)];
  if (j == 1 && r.plane[1][0]) v = v + r.plane[1][0][(i - 1) + (long)e0 * (k - 1)];
  if (j == e1 && r.plane[1][1]) v = v + r.plane[1][1][(i - 1) + (long)e0 * (k - 1)];
  if (k == 1 && r.plane[2][0]) v = v + r.plane[2][0][(i - 1) + (long)e0 * (j - 1)];
  if (k == r.ext[2] && r.plane[2][1]) v = v + r.plane[2][1][(i - 1) + (long)e0 * (j - 1)];
  out[c] = v;
}

// p += pp, then p += alpha * (Laplacian of pp) (the reference's two loops)
__global__ __launch_bounds__(TX *TY) void k_updt_pressure_imp(Geo g, Metrics m, const double *__restrict__ pp,
                                                              double *__restrict__ p, double alpha) {
  const int i = blockIdx.x * TX + threadIdx.x + 1;
  const int j = blockIdx.y * TY + threadIdx.y + 1;
  const int k = blockIdx.z + 1;
  if (i > g.nx || j > g.ny) return;
  const long sx = g.sx, sz = g.sx * g.sy;
  const long c = i + sx * j + sz * (long)k;
  const double pc = pp[c];
  double pv = p[c] + pc;
  pv = pv + alpha * (((pp[c + 1] - pc) * m.idxc[i] - (pc - pp[c - 1]) * m.idxc[i - 1]) * m.idxf[i] +
                     ((pp[c + sx] - pc) * m.idyc[j] - (pc - pp[c - sx]) * m.idyc[j - 1]) * m.idyf[j] +
                     ((pp[c + sz] - pc) * m.idzc[k] - (pc - pp[c - sz]) * m.idzc[k - 1]) * m.idzf[k]);
  p[c] = pv;
}

// dst = src over the unknowns of a field (1..ext per axis)
__global__ __launch_bounds__(TX *TY) void k_copy_unknowns(Geo g, int e0, int e1, int e2, const double *__restrict__ src,
                                                          double *__restrict__ dst) {
  const int i = blockIdx.x * TX + threadIdx.x + 1;
  const int j = blockIdx.y * TY + threadIdx.y + 1;
  const int k = blockIdx.z + 1;
  if (i > e0 || j > e1 || k > e2) return;
  const long c = i + g.sx * (j + g.sy * (long)k);
  dst[c] = src[c];
}

dim3 interior_grid(const Geo &g) { return dim3((g.nx + TX - 1) / TX, (g.ny + TY - 1) / TY, g.nz); }

}  // namespace

void launch_imp_rhs(hipStream_t s, const Geo &g, const ImpRhs &r, const double *phi, double *out, double alphai) {
  hipLaunchKernelGGL(k_imp_rhs, interior_grid(g), dim3(TX, TY), 0, s, g, r, phi, out, alphai);
  SNAC_HIP(hipGetLastError());
}

void launch_copy_unknowns(hipStream_t s, const Geo &g, const int ext[3], const double *src, double *dst) {
  hipLaunchKernelGGL(k_copy_unknowns, interior_grid(g), dim3(TX, TY), 0, s, g, ext[0], ext[1], ext[2], src, dst);
  SNAC_HIP(hipGetLastError());
}

void launch_updt_pressure_imp(hipStream_t s, const Geo &g, const Metrics &m, const double *pp, double *p,
                              double alpha) {
  hipLaunchKernelGGL(k_updt_pressure_imp, interior_grid(g), dim3(TX, TY), 0, s, g, m, pp, p, alpha);
  SNAC_HIP(hipGetLastError());
}

}  // namespace snac
