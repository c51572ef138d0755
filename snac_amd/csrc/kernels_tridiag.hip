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

// 1/x from the hardware estimate refined by two Newton steps (within a few
// ulp of the IEEE quotient): the eliminations' reciprocals without the IEEE
// division sequence on their dependent chains
__device__ __forceinline__ double rcp_ref(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = __builtin_fma(-x, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-x, r, 1.0);
  return __builtin_fma(r, e, r);
}

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
  const double lam = t.lamA[t1] + t.lamB[t2] + t.shift;
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

// Chunked variant without scratch arrays: a workgroup owns 64 lines (one
// per lane) and NC wavefronts, wavefront w the w-th chunk of every line's
// rows (at most K rows).  Each lane eliminates its chunk's interior
// (downwards, then upwards), leaving two rows per chunk that couple only the
// chunk's first and last unknowns with the neighbouring chunks'; those 2*NC
// rows form a tridiagonal system per line, solved in LDS by the first
// wavefront; every wavefront then recovers its interior rows.  HBM traffic:
// the rhs read once, the solution written once.  Lines with lam == 0 and
// pin_when_zero take x = 0 on their last row; a zero pivot (the
// all-periodic zero mode) gives 0.
template <int NC, int K>
__global__ __launch_bounds__(64 * NC) void k_tridiag_chunk(TridiagLine t, double *__restrict__ x, int base, int extra) {
  constexpr int NR = 2 * NC;
  __shared__ double red[4 * NR * 64];  // [comp][row][lane]: a, b, c (then c'), d (then d', then x)
  // the chunk index is wave-uniform: scalar registers, scalar coefficient loads
  const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int s = t.s;
  const int o1 = s == 0 ? 1 : 0, o2 = s == 2 ? 1 : 2;
  const long st[3] = {1, (long)t.G[0], (long)t.G[0] * t.G[1]};
  const int n1 = t.G[o1], n2 = t.G[o2], m = t.G[s];
  const long line = (long)blockIdx.x * 64 + lane;
  const bool live = line < (long)n1 * n2;
  const int t1 = live ? (int)(line % n1) : 0, t2 = live ? (int)(line / n1) : 0;
  const double lam = t.lamA[t1] + t.lamB[t2] + t.shift;
  const bool pin = t.pin_when_zero && lam == 0.0;
  const long ss = st[s];
  double *xl = x + t1 * st[o1] + t2 * st[o2];
  const int cnt = base + (w < extra ? 1 : 0);
  const int p0 = w * base + min(w, extra);
  // row q: lower l, diagonal d, upper u (the pinned row is the identity)
  auto coef = [&](int q, double &l, double &d, double &u) {
    l = q > 0 ? t.am[q] : 0.0;
    u = q < m - 1 ? t.ap[q] : 0.0;
    d = t.d0[q] + lam * t.wt[q];
    if (pin && q == m - 1) { l = 0.0; d = 1.0; u = 0.0; }
  };
  double A[K], C[K], D[K];
#pragma unroll
  for (int j = 0; j < K; j++) {
    A[j] = 0.0; C[j] = 0.0;
    const int q = p0 + j;
    D[j] = (j < cnt && live && !(pin && q == m - 1)) ? xl[q * ss] : 0.0;
  }
  // downward: row j >= 1 becomes A x_p0 + x_j + C x_{j+1} = D
  double aq = 0.0, cq = 0.0, dq = 0.0;
#pragma unroll
  for (int j = 1; j < K; j++) {
    if (j < cnt) {
      double l, d, u;
      coef(p0 + j, l, d, u);
      const double den = (j == 1) ? d : d - l * C[j - 1];
      const double ir = den != 0.0 ? rcp_ref(den) : 0.0;
      if (j == 1) { A[1] = l * ir; C[1] = u * ir; D[1] = D[1] * ir; }
      else { A[j] = -l * A[j - 1] * ir; C[j] = u * ir; D[j] = (D[j] - l * D[j - 1]) * ir; }
      if (j == cnt - 1) { aq = A[j]; cq = C[j]; dq = D[j]; }
    }
  }
  // upward: x_j = D - A x_p0 - C x_last for 1 <= j <= cnt-2
  double dn = 0.0, an = 0.0, cn = 0.0;
#pragma unroll
  for (int j = K - 2; j >= 1; j--) {
    if (j <= cnt - 2) {
      if (j == cnt - 2) { dn = D[j]; an = A[j]; cn = C[j]; }
      else { dn = D[j] - C[j] * dn; an = A[j] - C[j] * an; cn = -C[j] * cn; }
      D[j] = dn; A[j] = an; C[j] = cn;
    }
  }
  // the chunk's two rows of the reduced system (u[2w] = x_first, u[2w+1] = x_last)
  {
    double l0, d0, u0;
    coef(p0, l0, d0, u0);
    double r0b, r0g, r0d;
    if (cnt == 2) { r0b = d0; r0g = u0; r0d = D[0]; }
    else { r0b = d0 - u0 * an; r0g = -u0 * cn; r0d = D[0] - u0 * dn; }
    const int k0 = 2 * w, k1 = 2 * w + 1;
    red[(0 * NR + k0) * 64 + lane] = l0; red[(1 * NR + k0) * 64 + lane] = r0b;
    red[(2 * NR + k0) * 64 + lane] = r0g; red[(3 * NR + k0) * 64 + lane] = r0d;
    red[(0 * NR + k1) * 64 + lane] = aq; red[(1 * NR + k1) * 64 + lane] = 1.0;
    red[(2 * NR + k1) * 64 + lane] = cq; red[(3 * NR + k1) * 64 + lane] = dq;
  }
  __syncthreads();
  if (w == 0) {  // Thomas on this lane's 2*NC reduced rows, c' and d' kept in LDS
    double cprev = 0.0, dprev = 0.0;
#pragma unroll 1
    for (int k = 0; k < NR; k++) {
      const double a = red[(0 * NR + k) * 64 + lane], b = red[(1 * NR + k) * 64 + lane];
      const double g = red[(2 * NR + k) * 64 + lane], d = red[(3 * NR + k) * 64 + lane];
      const double den = k == 0 ? b : b - a * cprev;
      const double ir = den != 0.0 ? rcp_ref(den) : 0.0;
      cprev = g * ir;
      dprev = (d - (k == 0 ? 0.0 : a * dprev)) * ir;
      red[(2 * NR + k) * 64 + lane] = cprev;
      red[(3 * NR + k) * 64 + lane] = dprev;
    }
    double xv = dprev;
#pragma unroll 1
    for (int k = NR - 2; k >= 0; k--) {
      xv = red[(3 * NR + k) * 64 + lane] - red[(2 * NR + k) * 64 + lane] * xv;
      red[(3 * NR + k) * 64 + lane] = xv;
    }
  }
  __syncthreads();
  if (!live) return;
  const double xp = red[(3 * NR + 2 * w) * 64 + lane], xq = red[(3 * NR + 2 * w + 1) * 64 + lane];
  xl[p0 * ss] = xp;
  xl[(p0 + cnt - 1) * ss] = xq;
#pragma unroll
  for (int j = 1; j < K - 1; j++)
    if (j <= cnt - 2) xl[(p0 + j) * ss] = D[j] - A[j] * xp - C[j] * xq;
}

}  // namespace

// chunk count for the chunked variant (0: not applicable to this length):
// at most 16 rows per chunk, at least 2; fine: twice that many chunks (at
// most 16, at least 2 rows each), for more lines in flight per CU
int tridiag_chunks(int m, bool fine) {
  for (int nc : {4, 8, 16})
    if (m >= 2 * nc && (m + nc - 1) / nc <= 16) {
      if (!fine) return nc;
      return 2 * nc <= 16 && m >= 4 * nc ? 2 * nc : 0;
    }
  return 0;
}

void launch_tridiag(hipStream_t st, const TridiagLine &t, double *x, double *dscratch, double *cscratch) {
  const int s = t.s;
  const int o1 = s == 0 ? 1 : 0, o2 = s == 2 ? 1 : 2;
  long nlines = (long)t.G[o1] * t.G[o2];
  dim3 grid((unsigned)((nlines + WG - 1) / WG));
  const int nc = t.variant == 2 ? tridiag_chunks(t.G[s]) : (t.variant == 3 ? tridiag_chunks(t.G[s], true) : 0);
  if (nc > 0) {
    const int base = t.G[s] / nc, extra = t.G[s] % nc;
    const bool small = base + (extra ? 1 : 0) <= 9;
    if (nc == 4) {
      if (small) hipLaunchKernelGGL((k_tridiag_chunk<4, 9>), grid, dim3(256), 0, st, t, x, base, extra);
      else hipLaunchKernelGGL((k_tridiag_chunk<4, 17>), grid, dim3(256), 0, st, t, x, base, extra);
    } else if (nc == 8) {
      if (small) hipLaunchKernelGGL((k_tridiag_chunk<8, 9>), grid, dim3(512), 0, st, t, x, base, extra);
      else hipLaunchKernelGGL((k_tridiag_chunk<8, 17>), grid, dim3(512), 0, st, t, x, base, extra);
    } else {
      if (small) hipLaunchKernelGGL((k_tridiag_chunk<16, 9>), grid, dim3(1024), 0, st, t, x, base, extra);
      else hipLaunchKernelGGL((k_tridiag_chunk<16, 17>), grid, dim3(1024), 0, st, t, x, base, extra);
    }
  } else if (t.G[s] <= LDS_M && t.variant == 0)
    hipLaunchKernelGGL(k_tridiag<true>, grid, dim3(WG), 0, st, t, x, dscratch, cscratch);
  else
    hipLaunchKernelGGL(k_tridiag<false>, grid, dim3(WG), 0, st, t, x, dscratch, cscratch);
  SNAC_HIP(hipGetLastError());
}

}  // namespace snac
