// Tridiagonal line solves of the tensor-product Poisson path: after the
// r2r transforms along every uniform axis, each remaining line along the
// (possibly stretched) axis s is one symmetric tridiagonal system whose
// diagonal carries the eigenvalues of the transformed axes (fft.f90:151-189)
// times the cell weight, i.e. the per-mode 2D matrices of init_matrix_2d
// (solver.f90:687-842) plus lambda on the diagonal (init_n_2d_matrices
// solver.f90:981-1020), or the 3D matrix of init_matrix_3d (158-372),
// diagonalised along the transformed axes.  One thread per line (Thomas
// algorithm); consecutive threads own neighbouring lines so each sweep step
// reads a contiguous row.
#include "common.hpp"

namespace snac {
namespace {

constexpr int WG = 64;
constexpr int LDS_M = 128;  // line length served from LDS (64 lines x 128 x 8 B)

// Thomas algorithm, one thread per line.  The forward sweep reads the rhs
// from x and writes d' to `dsc`, a separate array, so the loads of the next
// U rows can be issued ahead of the dependent divisions; c' lives in LDS
// (or in `csc` for lines longer than LDS_M).  The backward sweep reads d'
// and c' and writes the solution back into x.
template <bool USE_LDS>
__global__ __launch_bounds__(WG) void k_tridiag(TridiagLine t, double *__restrict__ x, double *__restrict__ dsc,
                                                 double *__restrict__ csc) {
  constexpr int U = 8;
  __shared__ double cps[USE_LDS ? WG * LDS_M : 1];
  const int s = t.s;
  const int o1 = s == 0 ? 1 : 0, o2 = s == 2 ? 1 : 2;
  const long st[3] = {1, (long)t.G[0], (long)t.G[0] * t.G[1]};
  const int n1 = t.G[o1], n2 = t.G[o2], m = t.G[s];
  const long line = (long)blockIdx.x * WG + threadIdx.x;
  if (line >= (long)n1 * n2) return;
  const int t1 = (int)(line % n1), t2 = (int)(line / n1);
  const double lam = t.lamA[t1] + t.lamB[t2];
  const int mm = (t.pin_when_zero && lam == 0.0) ? m - 1 : m;
  const long base = t1 * st[o1] + t2 * st[o2];
  const long ss = st[s];
  double *xl = x + base;
  double *dl = dsc + base;
  double *cp = USE_LDS ? cps + threadIdx.x : csc + base;
  const long cs = USE_LDS ? WG : ss;
  double cprev = 0.0, dprev = 0.0;
  for (int q0 = 0; q0 < mm; q0 += U) {
    double xv[U];
#pragma unroll
    for (int u = 0; u < U; u++) xv[u] = (q0 + u < mm) ? xl[(q0 + u) * ss] : 0.0;
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int q = q0 + u;
      if (q < mm) {
        const double a = t.am[q], c = t.ap[q];
        const double b = t.d0[q] + lam * t.wt[q];
        const double den = (q == 0) ? b : b - a * cprev;
        // a zero pivot only occurs for the singular zero mode of an
        // all-periodic problem (diagonal system): its unknown is set to 0
        const double cq = den != 0.0 ? c / den : 0.0;
        const double dq = den != 0.0 ? (xv[u] - (q == 0 ? 0.0 : a * dprev)) / den : 0.0;
        cp[q * cs] = cq;
        dl[q * ss] = dq;
        cprev = cq;
        dprev = dq;
      }
    }
  }
  if (mm < m) xl[(m - 1) * ss] = 0.0;
  if (mm == 0) return;
  double xv = dprev;
  xl[(mm - 1) * ss] = xv;
  for (int q0 = mm - 2; q0 >= 0; q0 -= U) {
    double dv[U], cv[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int q = q0 - u;
      dv[u] = (q >= 0) ? dl[q * ss] : 0.0;
      cv[u] = (q >= 0) ? cp[q * cs] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int q = q0 - u;
      if (q >= 0) {
        xv = dv[u] - cv[u] * xv;
        xl[q * ss] = xv;
      }
    }
  }
}

}  // namespace

void launch_tridiag(hipStream_t st, const TridiagLine &t, double *x, double *dscratch, double *cscratch) {
  const int s = t.s;
  const int o1 = s == 0 ? 1 : 0, o2 = s == 2 ? 1 : 2;
  long nlines = (long)t.G[o1] * t.G[o2];
  dim3 grid((unsigned)((nlines + WG - 1) / WG));
  if (t.G[s] <= LDS_M && t.variant == 0)
    hipLaunchKernelGGL(k_tridiag<true>, grid, dim3(WG), 0, st, t, x, dscratch, cscratch);
  else
    hipLaunchKernelGGL(k_tridiag<false>, grid, dim3(WG), 0, st, t, x, dscratch, cscratch);
  SNAC_HIP(hipGetLastError());
}

}  // namespace snac
