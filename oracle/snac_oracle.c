/*
 * snac_oracle.c -- CPU FP64 restatement of SNaC's time-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker for the HIP path; never shipped, never
 * called by the product library).  Paths cited are relative to the SNaC tree.
 *
 * Parity pins (each checked by tests/):
 *   grid (initgrid.f90)           : closed form (uniform spacing, tanh map),
 *                                   reference LDC golden end-to-end
 *   block placement (initmpi.f90) : index fixtures for the reference's own
 *                                   multi-block examples, LDC golden
 *   rk_mom / mom_*                : closed-form checks (linear field has zero
 *                                   diffusion, periodic integral of the
 *                                   vv-advection is zero), LDC golden
 *   bounduvw/boundp/inflow        : ghost-value fixtures on a tiny multi-block
 *                                   case, LDC golden
 *   fillps/correc/updt_pressure   : discrete div-free field gives zero rhs,
 *                                   projection gives div ~ 0, LDC golden
 *   r2r kinds                     : scipy.fft dct/dst (types II,III,IV) and
 *                                   numpy rfft fixtures; direct vs fast path
 *   eigenvalues                   : numpy eigvalsh of the 1D operator
 *   operator assembly + solvers   : scipy sparse direct solve; residual check
 *   chkdiv / chkdt                : closed-form fields
 *   implicit diffusion (_IMPDIFF) : Helmholtz solves of u, v, w on discrete
 *                                   eigenmodes of each staggered operator
 *                                   (periodic, centred/staggered Dirichlet,
 *                                   staggered Neumann; FFT and 3D forms),
 *                                   constants / linear fields under
 *                                   Dirichlet / Neumann values, the per-step
 *                                   damping of velocity and scalar waves,
 *                                   the updt_pressure Laplacian term.  The
 *                                   inflow-array scaling of init_bc_rhs is
 *                                   checked only on a uniform inflow
 *                                   profile.
 * The HYPRE iterates and FFTW rounding of the reference cannot be reproduced
 * bit-wise (third-party, not vendored); the oracle solves the same assembled
 * systems to a tight tolerance instead.
 */
#include "snac_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MAXB 64
#define MAXS 8
#define PI 3.14159265358979323846264338327950288

typedef struct {
  int n[3], lo_g[3], nb[2][3], is_bound[2][3];
  double lmin[3], lmax[3], gr[3];
  int gt[3];
  char cbcvel[3][3][2];  /* [comp][dir][side] */
  double bcvel[3][3][2];
  char cbcpre[3][2];
  double bcpre[3][2];
  int inflow_type[3][2], is_inflow[3][2];
  double *dc[3], *df[3], *rc[3], *rf[3]; /* n+2 entries */
  double *f[6];                           /* u v w p pp po, ghosted */
  double *dto[3];                         /* dudtrko... interior */
  double *vin[3][3][2];                   /* inflow planes [idir][comp][side] */
  double *rhsp[3][2];                     /* BC rhs planes [dir][side] */
  double *sv[MAXS], *sdo[MAXS];           /* scalars (ghosted) and dsdtrko (interior) */
  char cbcsca[MAXS][3][2];                /* scalar BCs [iscal][dir][side] */
  double bcsca[MAXS][3][2];
  long sx, sy, ntot;                      /* strides of ghosted arrays */
} ob;

/* one elliptic operator (the pressure's, or the Helmholtz operator of a
 * velocity component or scalar with implicit diffusion)                     */
typedef struct {
  int which;       /* 0 pressure, 1..3 u, v, w, 4+i scalar i */
  int stag;        /* staggered axis of a velocity component, -1 otherwise */
  int nmodes, qfft, kind_fwd, kind_bwd;
  double *lambda, normfft;
  int nd, ax[3];   /* reduced problem dimensionality and its axes */
  int ext[MAXB][3];          /* unknowns of each block along x, y, z */
  long off[MAXB], nunk;      /* offsets of the blocks in the solver vector */
  double *diag0, *wgt, *coef; /* per unknown: base diagonal, lambda weight, nd*2 coefs */
  long *nbr;                   /* per unknown nd*2 neighbour indices (-1 none) */
  int has_dirichlet;
  double shift;                /* added to lambda: alphai of a Helmholtz solve */
  /* band factorizations, lazily built per mode, kept for three shifts */
  long *perm, *iperm; long bw;
  double **bandL[3]; double band_shift[3]; int band_next;
  int band_ok;
  /* tensor solver data */
  int tensor_ok, tt[2], nt, ts; double th[2];
  int t_kind_fwd[2], t_kind_bwd[2]; double t_norm[2]; double *t_lambda[2];
  double *work;
  double *rhs[MAXB][3][2];     /* init_bc_rhs planes over the unknowns (Helmholtz) */
} oop;

struct oc_case {
  int nb, fft_axis;
  double visc, bforce[3], vol_all;
  /* scalars (param.f90:96-100, scal.f90:22-49) */
  int nscal;
  double salpha[MAXS], ssource[MAXS], scalf[MAXS], sf[MAXS], beta, gacc[3], alpha_max;
  int buoyancy; /* the _BOUSSINESQ_BUOYANCY build */
  int is_sforced[MAXS];
  int periods[3], is_periodic[3];
  ob b[MAXB];
  /* solver */
  int method, maxiter;
  double tol;
  oop P;                 /* pressure */
  int impdiff;           /* implicit diffusion (the reference's _IMPDIFF build) */
  oop H[3 + MAXS];       /* Helmholtz operators of u, v, w and the scalars */
};

static int nthreads_set = 0;
int oc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) { omp_set_num_threads(n); nthreads_set = n; }
  return omp_get_max_threads();
#else
  (void)n; return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* grid: initgrid.f90:9-83 and the stretching functions 171-287             */
/* ------------------------------------------------------------------------ */
static double gp_geometric_one_end(double alpha, double r0) {
  const double power = 1.0 / 3.0;
  if (r0 == 1.0) return r0;
  return r0 * pow((1.0 - pow(r0, alpha)) / (1.0 - r0) / alpha, power);
}
static double gridpoint(int gt, double a, double r0) {
  switch (gt) {
  default:
  case 0: /* two ends, initgrid.f90:171-184 */
    return (a != 0.0) ? 0.5 * (1.0 + tanh((r0 - 0.5) * a) / tanh(a / 2.0)) : r0;
  case 1: /* one end (lower) */
    return (a != 0.0) ? 1.0 * (1.0 + tanh((r0 - 1.0) * a) / tanh(a / 1.0)) : r0;
  case 2: /* middle */
    if (a == 0.0) return r0;
    if (r0 <= 0.5) return 0.5 * (1.0 - 1.0 + tanh(2.0 * a * (r0 - 0.0)) / tanh(a));
    return 0.5 * (1.0 + 1.0 + tanh(2.0 * a * (r0 - 1.0)) / tanh(a));
  case 3: /* one end (upper) */
    return (a != 0.0) ? 1.0 - 1.0 * (1.0 + tanh((1.0 - r0 - 1.0) * a) / tanh(a / 1.0)) : r0;
  case 4: return gp_geometric_one_end(a, r0);
  case 5: return 1.0 - gp_geometric_one_end(a, 1.0 - r0);
  case 6:
    if (r0 <= 0.5) return gp_geometric_one_end(a / 2.0, 2.0 * r0) / 2.0;
    return 1.0 - gp_geometric_one_end(a / 2.0, 2.0 * (1.0 - r0)) / 2.0;
  case 7:
    if (r0 <= 0.5) return (1.0 - gp_geometric_one_end(a / 2.0, (0.5 - r0) * 2.0)) / 2.0;
    return (1.0 + gp_geometric_one_end(a / 2.0, (r0 - 0.5) * 2.0)) / 2.0;
  }
}

int oc_initgrid(int n, int gt, double gr, double lmin, double lmax,
                double *drc, double *drf, double *rc, double *rf) {
  if (n < 1) return -1;
  for (int q = 1; q <= n; q++) {
    double r0 = (double)q / (double)n;
    rf[q] = lmin + gridpoint(gt, gr, r0) * (lmax - lmin);
  }
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
  return 0;
}

/* ------------------------------------------------------------------------ */
/* field helpers                                                             */
/* ------------------------------------------------------------------------ */
#define IDX(B, i, j, k) ((long)(i) + (B)->sx * ((long)(j) + (B)->sy * (long)(k)))
#define IIDX(B, i, j, k) ((long)((i)-1) + (long)(B)->n[0] * ((long)((j)-1) + (long)(B)->n[1] * (long)((k)-1)))

static void *xcalloc(size_t n, size_t s) {
  void *p = calloc(n ? n : 1, s);
  if (!p) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
  return p;
}

/* plane of a ghosted field: dir d, plane index q (0..n+1), element (t1,t2)
 * over the full ghosted tangential ranges (t1 along the lower remaining axis). */
static inline long plane_idx(const ob *B, int d, int q, int t1, int t2) {
  switch (d) {
  case 0: return IDX(B, q, t1, t2);
  case 1: return IDX(B, t1, q, t2);
  default: return IDX(B, t1, t2, q);
  }
}
static inline void tang(int d, int *a1, int *a2) {
  if (d == 0) { *a1 = 1; *a2 = 2; }
  else if (d == 1) { *a1 = 0; *a2 = 2; }
  else { *a1 = 0; *a2 = 1; }
}

/* ------------------------------------------------------------------------ */
/* boundary conditions: bound.f90                                            */
/* ------------------------------------------------------------------------ */
/* set_bc bound.f90:108-248 on one ghost plane of one block                  */
static void set_bc(ob *B, double *p, char ctype, int ibound, int d, int centered,
                   double rvalue, double dr) {
  double factor = rvalue, sgn = 0.0;
  if (ctype == 'D' && centered) { factor = 2.0 * factor; sgn = -1.0; }
  if (ctype == 'N') {
    if (ibound == 0) factor = -dr * factor; else factor = dr * factor;
    sgn = 1.0;
  }
  if (ctype != 'D' && ctype != 'N') return;
  int a1, a2; tang(d, &a1, &a2);
  int m1 = B->n[a1] + 2, m2 = B->n[a2] + 2, n = B->n[d];
  for (int t2 = 0; t2 < m2; t2++)
    for (int t1 = 0; t1 < m1; t1++) {
      if (centered) {
        if (ibound == 0) p[plane_idx(B, d, 0, t1, t2)] = factor + sgn * p[plane_idx(B, d, 1, t1, t2)];
        else p[plane_idx(B, d, n + 1, t1, t2)] = factor + sgn * p[plane_idx(B, d, n, t1, t2)];
      } else if (ctype == 'D') {
        if (ibound == 0) p[plane_idx(B, d, 0, t1, t2)] = factor;
        else {
          p[plane_idx(B, d, n, t1, t2)] = factor;
          p[plane_idx(B, d, n + 1, t1, t2)] = p[plane_idx(B, d, n - 1, t1, t2)];
        }
      } else { /* staggered N */
        if (ibound == 0) p[plane_idx(B, d, 0, t1, t2)] = factor + p[plane_idx(B, d, 1, t1, t2)];
        else {
          p[plane_idx(B, d, n, t1, t2)] = factor + p[plane_idx(B, d, n - 1, t1, t2)];
          p[plane_idx(B, d, n + 1, t1, t2)] = p[plane_idx(B, d, n, t1, t2)];
        }
      }
    }
}

/* updthalo bound.f90:587-650: with every block on one address space, the two
 * MPI_SENDRECVs of direction d become plane copies from the neighbour blocks'
 * first/last interior planes (full ghosted tangential extents, which match
 * between friend blocks, initmpi.f90:165-175).  Reads touch only interior
 * planes along d and writes only ghost planes along d, so the sweep over
 * blocks is order independent.                                               */
/* field id -> array: 0..5 u,v,w,p,pp,po ; 9+2i scalar i                    */
static double *farr(ob *B, int fid) {
  if (fid >= 0 && fid < 6) return B->f[fid];
  if (fid >= 9 && (fid - 9) % 2 == 0 && (fid - 9) / 2 < MAXS) return B->sv[(fid - 9) / 2];
  return NULL;
}

static void updthalo(oc_case *c, int fid, int d) {
  for (int ib = 0; ib < c->nb; ib++) {
    ob *B = &c->b[ib];
    int a1, a2; tang(d, &a1, &a2);
    int m1 = B->n[a1] + 2, m2 = B->n[a2] + 2;
    for (int side = 0; side < 2; side++) {
      int jb = B->nb[side][d];
      if (jb < 0) continue;
      ob *N = &c->b[jb];
      int qdst = side == 0 ? 0 : B->n[d] + 1;
      int qsrc = side == 0 ? N->n[d] : 1;
      double *dst = farr(B, fid);
      const double *src = farr(N, fid);
      for (int t2 = 0; t2 < m2; t2++)
        for (int t1 = 0; t1 < m1; t1++)
          dst[plane_idx(B, d, qdst, t1, t2)] = src[plane_idx(N, d, qsrc, t1, t2)];
    }
  }
}

/* bounduvw bound.f90:8-66                                                   */
void oc_bounduvw(oc_case *c, int is_correc) {
  for (int f = 0; f < 3; f++)
    for (int d = 0; d < 3; d++) updthalo(c, f, d);
  for (int ib = 0; ib < c->nb; ib++) {
    ob *B = &c->b[ib];
    for (int d = 0; d < 3; d++)
      for (int side = 0; side < 2; side++) {
        if (!B->is_bound[side][d]) continue;
        int qm = side == 0 ? 0 : B->n[d]; /* dr index: lo-1 or hi */
        for (int comp = 0; comp < 3; comp++) {
          int centered = comp != d;
          if (!centered && is_correc) continue;
          double dr = centered ? B->dc[d][qm] : B->df[d][qm];
          set_bc(B, B->f[comp], B->cbcvel[comp][d][side], side, d, centered,
                 B->bcvel[comp][d][side], dr);
        }
      }
  }
}

/* boundp bound.f90:68-106 ; for a scalar (which = 9 + 2*iscal) with the
 * scalar's own BCs, as main.f90:565-568,934 call it                         */
void oc_boundp(oc_case *c, int which) {
  int is = which >= 9 ? (which - 9) / 2 : -1;
  if (is >= c->nscal) return;
  for (int d = 0; d < 3; d++) updthalo(c, which, d);
  for (int ib = 0; ib < c->nb; ib++) {
    ob *B = &c->b[ib];
    for (int d = 0; d < 3; d++)
      for (int side = 0; side < 2; side++) {
        if (!B->is_bound[side][d]) continue;
        int qm = side == 0 ? 0 : B->n[d];
        if (is < 0)
          set_bc(B, B->f[which], B->cbcpre[d][side], side, d, 1, B->bcpre[d][side], B->dc[d][qm]);
        else
          set_bc(B, B->sv[is], B->cbcsca[is][d][side], side, d, 1, B->bcsca[is][d][side], B->dc[d][qm]);
      }
  }
}

/* inflow bound.f90:438-477                                                  */
void oc_inflow(oc_case *c, int is_correc) {
  for (int ib = 0; ib < c->nb; ib++) {
    ob *B = &c->b[ib];
    for (int d = 0; d < 3; d++)
      for (int side = 0; side < 2; side++) {
        if (!B->is_inflow[d][side]) continue;
        int n = B->n[d];
        int qn = side == 0 ? 0 : n;
        int qt = side == 0 ? 0 : n + 1;
        int dq = side == 0 ? 1 : -1;
        int a1, a2; tang(d, &a1, &a2);
        int m1 = B->n[a1] + 2, m2 = B->n[a2] + 2;
        for (int comp = 0; comp < 3; comp++) {
          double *F = B->f[comp];
          const double *vin = B->vin[d][comp][side];
          for (int t2 = 0; t2 < m2; t2++)
            for (int t1 = 0; t1 < m1; t1++) {
              double val = vin[t1 + (long)m1 * t2];
              if (comp == d) {
                if (!is_correc) F[plane_idx(B, d, qn, t1, t2)] = val;
              } else {
                F[plane_idx(B, d, qt, t1, t2)] = -F[plane_idx(B, d, qt + dq, t1, t2)] + 2.0 * val;
              }
            }
        }
      }
  }
}

/* ------------------------------------------------------------------------ */
/* momentum: rk.f90:7-128 with mom.f90:112-386                              */
/* ------------------------------------------------------------------------ */
static const double RKC[3][2] = {{32.0 / 60.0, 0.0}, {25.0 / 60.0, -17.0 / 60.0}, {45.0 / 60.0, -25.0 / 60.0}};

static void rk_mom_block(oc_case *c, ob *B, int irk, double dt) {
  const double f1 = RKC[irk - 1][0] * dt, f2 = RKC[irk - 1][1] * dt, f12 = f1 + f2;
  const double visc = c->visc;
  const int nx = B->n[0], ny = B->n[1], nz = B->n[2];
  const double *dxc = B->dc[0], *dxf = B->df[0], *dyc = B->dc[1], *dyf = B->df[1],
               *dzc = B->dc[2], *dzf = B->df[2];
  double *u = B->f[0], *v = B->f[1], *w = B->f[2];
  const double *p = B->f[3];
  long nin = (long)nx * ny * nz;
  double *du = xcalloc(nin, sizeof(double)), *dv = xcalloc(nin, sizeof(double)),
         *dw = xcalloc(nin, sizeof(double));
  /* _IMPDIFF: the diffusion goes to dudtrkd.. (rk.f90:29-31,50-59) */
  const int imp = c->impdiff;
  double *dud = imp ? xcalloc(nin, sizeof(double)) : NULL, *dvd = imp ? xcalloc(nin, sizeof(double)) : NULL,
         *dwd = imp ? xcalloc(nin, sizeof(double)) : NULL;
#define U(i, j, k) u[IDX(B, i, j, k)]
#define V(i, j, k) v[IDX(B, i, j, k)]
#define W(i, j, k) w[IDX(B, i, j, k)]
#define P(i, j, k) p[IDX(B, i, j, k)]
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) {
        long q = IIDX(B, i, j, k);
        double a = 0.0, b = 0.0, cc = 0.0;
        /* momx_d mom.f90:221-256 */
        {
          double dudxp = (U(i + 1, j, k) - U(i, j, k)) / dxf[i + 1];
          double dudxm = (U(i, j, k) - U(i - 1, j, k)) / dxf[i];
          double dudyp = (U(i, j + 1, k) - U(i, j, k)) / dyc[j];
          double dudym = (U(i, j, k) - U(i, j - 1, k)) / dyc[j - 1];
          double dudzp = (U(i, j, k + 1) - U(i, j, k)) / dzc[k];
          double dudzm = (U(i, j, k) - U(i, j, k - 1)) / dzc[k - 1];
          a = a + visc * (dudxp - dudxm) / dxc[i] + visc * (dudyp - dudym) / dyf[j] +
              visc * (dudzp - dudzm) / dzf[k];
        }
        /* momy_d mom.f90:257-292 */
        {
          double dvdxp = (V(i + 1, j, k) - V(i, j, k)) / dxc[i];
          double dvdxm = (V(i, j, k) - V(i - 1, j, k)) / dxc[i - 1];
          double dvdyp = (V(i, j + 1, k) - V(i, j, k)) / dyf[j + 1];
          double dvdym = (V(i, j, k) - V(i, j - 1, k)) / dyf[j];
          double dvdzp = (V(i, j, k + 1) - V(i, j, k)) / dzc[k];
          double dvdzm = (V(i, j, k) - V(i, j, k - 1)) / dzc[k - 1];
          b = b + visc * (dvdxp - dvdxm) / dxf[i] + visc * (dvdyp - dvdym) / dyc[j] +
              visc * (dvdzp - dvdzm) / dzf[k];
        }
        /* momz_d mom.f90:293-327 */
        {
          double dwdxp = (W(i + 1, j, k) - W(i, j, k)) / dxc[i];
          double dwdxm = (W(i, j, k) - W(i - 1, j, k)) / dxc[i - 1];
          double dwdyp = (W(i, j + 1, k) - W(i, j, k)) / dyc[j];
          double dwdym = (W(i, j, k) - W(i, j - 1, k)) / dyc[j - 1];
          double dwdzp = (W(i, j, k + 1) - W(i, j, k)) / dzf[k + 1];
          double dwdzm = (W(i, j, k) - W(i, j, k - 1)) / dzf[k];
          cc = cc + visc * (dwdxp - dwdxm) / dxf[i] + visc * (dwdyp - dwdym) / dyf[j] +
               visc * (dwdzp - dwdzm) / dzc[k];
        }
        if (imp) { dud[q] = a; dvd[q] = b; dwd[q] = cc; a = 0.0; b = 0.0; cc = 0.0; }
        /* momx_a_vv mom.f90:112-147 */
        {
          double uuip = 0.25 * (U(i, j, k) * dyf[j] * dzf[k] + U(i + 1, j, k) * dyf[j] * dzf[k]) * U(i + 1, j, k);
          double uuim = 0.25 * (U(i, j, k) * dyf[j] * dzf[k] + U(i - 1, j, k) * dyf[j] * dzf[k]) * U(i - 1, j, k);
          double uvjp = 0.25 * (V(i, j, k) * dxf[i] * dzf[k] + V(i + 1, j, k) * dxf[i + 1] * dzf[k]) * U(i, j + 1, k);
          double uvjm = 0.25 * (V(i, j - 1, k) * dxf[i] * dzf[k] + V(i + 1, j - 1, k) * dxf[i + 1] * dzf[k]) * U(i, j - 1, k);
          double uwkp = 0.25 * (W(i, j, k) * dxf[i] * dyf[j] + W(i + 1, j, k) * dxf[i + 1] * dyf[j]) * U(i, j, k + 1);
          double uwkm = 0.25 * (W(i, j, k - 1) * dxf[i] * dyf[j] + W(i + 1, j, k - 1) * dxf[i + 1] * dyf[j]) * U(i, j, k - 1);
          a = a + ((-uuip + uuim) + (-uvjp + uvjm) + (-uwkp + uwkm)) / (dxc[i] * dyf[j] * dzf[k]);
        }
        /* momy_a_vv mom.f90:148-183 */
        {
          double uvip = 0.25 * (U(i, j, k) * dyf[j] * dzf[k] + U(i, j + 1, k) * dyf[j + 1] * dzf[k]) * V(i + 1, j, k);
          double uvim = 0.25 * (U(i - 1, j, k) * dyf[j] * dzf[k] + U(i - 1, j + 1, k) * dyf[j + 1] * dzf[k]) * V(i - 1, j, k);
          double vvjp = 0.25 * (V(i, j, k) * dxf[i] * dzf[k] + V(i, j + 1, k) * dxf[i] * dzf[k]) * V(i, j + 1, k);
          double vvjm = 0.25 * (V(i, j, k) * dxf[i] * dzf[k] + V(i, j - 1, k) * dxf[i] * dzf[k]) * V(i, j - 1, k);
          double wvkp = 0.25 * (W(i, j, k) * dyf[j] * dxf[i] + W(i, j + 1, k) * dyf[j + 1] * dxf[i]) * V(i, j, k + 1);
          double wvkm = 0.25 * (W(i, j, k - 1) * dyf[j] * dxf[i] + W(i, j + 1, k - 1) * dyf[j + 1] * dxf[i]) * V(i, j, k - 1);
          b = b + ((-uvip + uvim) + (-vvjp + vvjm) + (-wvkp + wvkm)) / (dxf[i] * dyc[j] * dzf[k]);
        }
        /* momz_a_vv mom.f90:185-219 */
        {
          double uwip = 0.25 * (U(i, j, k) * dzf[k] * dyf[j] + U(i, j, k + 1) * dzf[k + 1] * dyf[j]) * W(i + 1, j, k);
          double uwim = 0.25 * (U(i - 1, j, k) * dzf[k] * dyf[j] + U(i - 1, j, k + 1) * dzf[k + 1] * dyf[j]) * W(i - 1, j, k);
          double vwjp = 0.25 * (V(i, j, k) * dzf[k] * dxf[i] + V(i, j, k + 1) * dzf[k + 1] * dxf[i]) * W(i, j + 1, k);
          double vwjm = 0.25 * (V(i, j - 1, k) * dzf[k] * dxf[i] + V(i, j - 1, k + 1) * dzf[k + 1] * dxf[i]) * W(i, j - 1, k);
          double wwkp = 0.25 * (W(i, j, k) * dxf[i] * dyf[j] + W(i, j, k + 1) * dxf[i] * dyf[j]) * W(i, j, k + 1);
          double wwkm = 0.25 * (W(i, j, k) * dxf[i] * dyf[j] + W(i, j, k - 1) * dxf[i] * dyf[j]) * W(i, j, k - 1);
          cc = cc + ((-uwip + uwim) + (-vwjp + vwjm) + (-wwkp + wwkm)) / (dxf[i] * dyf[j] * dzc[k]);
        }
        du[q] = a; dv[q] = b; dw[q] = cc;
      }
  /* explicit update rk.f90:68-85 */
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) {
        long q = IIDX(B, i, j, k), g = IDX(B, i, j, k);
        u[g] = u[g] + f1 * du[q] + f2 * B->dto[0][q];
        v[g] = v[g] + f1 * dv[q] + f2 * B->dto[1][q];
        w[g] = w[g] + f1 * dw[q] + f2 * B->dto[2][q];
        if (imp) { /* rk.f90:75-79 */
          u[g] = u[g] + f12 * dud[q];
          v[g] = v[g] + f12 * dvd[q];
          w[g] = w[g] + f12 * dwd[q];
        }
        B->dto[0][q] = du[q]; B->dto[1][q] = dv[q]; B->dto[2][q] = dw[q];
      }
  /* pressure gradient + body force mom.f90:328-386, update rk.f90:86-111 */
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) {
        long g = IDX(B, i, j, k);
        double px = -(P(i + 1, j, k) - P(i, j, k)) / dxc[i] + c->bforce[0];
        double py = -(P(i, j + 1, k) - P(i, j, k)) / dyc[j] + c->bforce[1];
        double pz = -(P(i, j, k + 1) - P(i, j, k)) / dzc[k] + c->bforce[2];
        u[g] = u[g] + f12 * px;
        v[g] = v[g] + f12 * py;
        w[g] = w[g] + f12 * pz;
        if (c->buoyancy && c->nscal > 0) { /* Boussinesq buoyancy rk.f90:104-108 */
          const double *S = B->sv[0];
          u[g] = u[g] - f12 * c->gacc[0] * c->beta * 0.5 * (S[IDX(B, i + 1, j, k)] + S[g]);
          v[g] = v[g] - f12 * c->gacc[1] * c->beta * 0.5 * (S[IDX(B, i, j + 1, k)] + S[g]);
          w[g] = w[g] - f12 * c->gacc[2] * c->beta * 0.5 * (S[IDX(B, i, j, k + 1)] + S[g]);
        }
      }
  if (imp) { /* rhs of the Helmholtz equation rk.f90:112-127 */
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 1; k <= nz; k++)
      for (int j = 1; j <= ny; j++)
        for (int i = 1; i <= nx; i++) {
          long q = IIDX(B, i, j, k), g = IDX(B, i, j, k);
          u[g] = u[g] - .5 * f12 * dud[q];
          v[g] = v[g] - .5 * f12 * dvd[q];
          w[g] = w[g] - .5 * f12 * dwd[q];
        }
  }
#undef U
#undef V
#undef W
#undef P
  free(du); free(dv); free(dw); free(dud); free(dvd); free(dwd);
}

void oc_rk_mom(oc_case *c, int irk, double dt) {
  for (int ib = 0; ib < c->nb; ib++) rk_mom_block(c, &c->b[ib], irk, dt);
}

/* ------------------------------------------------------------------------ */
/* fillps.f90:7-54 + updt_rhs bound.f90:543-585                              */
/* ------------------------------------------------------------------------ */
void oc_fillps(oc_case *c, double dtrk) {
  for (int ib = 0; ib < c->nb; ib++) {
    ob *B = &c->b[ib];
    const double *dxf = B->df[0], *dyf = B->df[1], *dzf = B->df[2];
    const double *u = B->f[0], *v = B->f[1], *w = B->f[2];
    double *pp = B->f[4];
    int nx = B->n[0], ny = B->n[1], nz = B->n[2];
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 1; k <= nz; k++)
      for (int j = 1; j <= ny; j++)
        for (int i = 1; i <= nx; i++) {
          double dw = w[IDX(B, i, j, k)] - w[IDX(B, i, j, k - 1)];
          double dv = v[IDX(B, i, j, k)] - v[IDX(B, i, j - 1, k)];
          double du = u[IDX(B, i, j, k)] - u[IDX(B, i - 1, j, k)];
          double s;
          switch (c->fft_axis) {
          case 3: s = dw / dzf[k] * dxf[i] * dyf[j] + dv * dxf[i] + du * dyf[j]; break;
          case 2: s = dw * dxf[i] + dv / dyf[j] * dxf[i] * dzf[k] + du * dzf[k]; break;
          case 1: s = dw * dyf[j] + dv * dzf[k] + du / dxf[i] * dyf[j] * dzf[k]; break;
          default: s = dw * dxf[i] * dyf[j] + dv * dxf[i] * dzf[k] + du * dyf[j] * dzf[k]; break;
          }
          pp[IDX(B, i, j, k)] = s / dtrk;
        }
    /* updt_rhs: x, y, z boundary planes in that order */
    for (int d = 0; d < 3; d++)
      for (int side = 0; side < 2; side++) {
        if (!B->is_bound[side][d] || !B->rhsp[d][side]) continue;
        int a1, a2; tang(d, &a1, &a2);
        int q = side == 0 ? 1 : B->n[d];
        for (int t2 = 1; t2 <= B->n[a2]; t2++)
          for (int t1 = 1; t1 <= B->n[a1]; t1++)
            pp[plane_idx(B, d, q, t1, t2)] += B->rhsp[d][side][(t1 - 1) + (long)B->n[a1] * (t2 - 1)];
      }
  }
}

/* correc.f90:7-48                                                           */
void oc_correc(oc_case *c, double dtrk) {
  for (int ib = 0; ib < c->nb; ib++) {
    ob *B = &c->b[ib];
    const double *p = B->f[4];
    double *u = B->f[0], *v = B->f[1], *w = B->f[2];
    int nx = B->n[0], ny = B->n[1], nz = B->n[2];
    for (int k = 0; k <= nz + 1; k++)
      for (int j = 0; j <= ny + 1; j++)
        for (int i = 0; i <= nx; i++)
          u[IDX(B, i, j, k)] = u[IDX(B, i, j, k)] - dtrk * (p[IDX(B, i + 1, j, k)] - p[IDX(B, i, j, k)]) / B->dc[0][i];
    for (int k = 0; k <= nz + 1; k++)
      for (int j = 0; j <= ny; j++)
        for (int i = 0; i <= nx + 1; i++)
          v[IDX(B, i, j, k)] = v[IDX(B, i, j, k)] - dtrk * (p[IDX(B, i, j + 1, k)] - p[IDX(B, i, j, k)]) / B->dc[1][j];
    for (int k = 0; k <= nz; k++)
      for (int j = 0; j <= ny + 1; j++)
        for (int i = 0; i <= nx + 1; i++)
          w[IDX(B, i, j, k)] = w[IDX(B, i, j, k)] - dtrk * (p[IDX(B, i, j, k + 1)] - p[IDX(B, i, j, k)]) / B->dc[2][k];
  }
}

/* updt_pressure.f90:23-29 (no _IMPDIFF term)                                */
void oc_updt_pressure(oc_case *c) {
  for (int ib = 0; ib < c->nb; ib++) {
    ob *B = &c->b[ib];
    for (int k = 1; k <= B->n[2]; k++)
      for (int j = 1; j <= B->n[1]; j++)
        for (int i = 1; i <= B->n[0]; i++) B->f[3][IDX(B, i, j, k)] += B->f[4][IDX(B, i, j, k)];
  }
}

/* updt_pressure.f90:23-45 of the _IMPDIFF build: p += pp, then p += alpha
 * times the Laplacian of pp (alpha = -visc*dtrk/2, main.f90:935,1031)       */
void oc_updt_pressure_impdiff(oc_case *c, double alpha) {
  oc_updt_pressure(c);
  for (int ib = 0; ib < c->nb; ib++) {
    ob *B = &c->b[ib];
    const double *pp = B->f[4];
    double *p = B->f[3];
    const double *dxc = B->dc[0], *dxf = B->df[0], *dyc = B->dc[1], *dyf = B->df[1],
                 *dzc = B->dc[2], *dzf = B->df[2];
    for (int k = 1; k <= B->n[2]; k++)
      for (int j = 1; j <= B->n[1]; j++)
        for (int i = 1; i <= B->n[0]; i++) {
#define PP(i, j, k) pp[IDX(B, i, j, k)]
          p[IDX(B, i, j, k)] = p[IDX(B, i, j, k)] +
                               alpha * (((PP(i + 1, j, k) - PP(i, j, k)) / dxc[i] -
                                         (PP(i, j, k) - PP(i - 1, j, k)) / dxc[i - 1]) / dxf[i] +
                                        ((PP(i, j + 1, k) - PP(i, j, k)) / dyc[j] -
                                         (PP(i, j, k) - PP(i, j - 1, k)) / dyc[j - 1]) / dyf[j] +
                                        ((PP(i, j, k + 1) - PP(i, j, k)) / dzc[k] -
                                         (PP(i, j, k) - PP(i, j, k - 1)) / dzc[k - 1]) / dzf[k]);
#undef PP
        }
  }
}

/* chkdiv.f90:8-44                                                           */
void oc_chkdiv(const oc_case *c, double *divtot, double *divmax) {
  double tot = 0.0, mx = 0.0;
  for (int ib = 0; ib < c->nb; ib++) {
    const ob *B = &c->b[ib];
    const double *u = B->f[0], *v = B->f[1], *w = B->f[2];
    for (int k = 1; k <= B->n[2]; k++)
      for (int j = 1; j <= B->n[1]; j++)
        for (int i = 1; i <= B->n[0]; i++) {
          double div = (w[IDX(B, i, j, k)] - w[IDX(B, i, j, k - 1)]) / B->df[2][k] +
                       (v[IDX(B, i, j, k)] - v[IDX(B, i, j - 1, k)]) / B->df[1][j] +
                       (u[IDX(B, i, j, k)] - u[IDX(B, i - 1, j, k)]) / B->df[0][i];
          if (fabs(div) > mx) mx = fabs(div);
          tot = tot + div * B->df[0][i] * B->df[1][j] * B->df[2][k] / c->vol_all;
        }
  }
  *divtot = tot;
  *divmax = mx;
}

/* chkdt.f90:8-56; the diffusive limit only without implicit diffusion      */
double oc_chkdt(const oc_case *c) {
  double dti = 0.0, dlmin = 1e300;
  for (int ib = 0; ib < c->nb; ib++) {
    const ob *B = &c->b[ib];
    const double *u = B->f[0], *v = B->f[1], *w = B->f[2];
    const double *dxc = B->dc[0], *dxf = B->df[0], *dyc = B->dc[1], *dyf = B->df[1],
                 *dzc = B->dc[2], *dzf = B->df[2];
    for (int k = 1; k <= B->n[2]; k++)
      for (int j = 1; j <= B->n[1]; j++)
        for (int i = 1; i <= B->n[0]; i++) {
#define U(i, j, k) u[IDX(B, i, j, k)]
#define V(i, j, k) v[IDX(B, i, j, k)]
#define W(i, j, k) w[IDX(B, i, j, k)]
          double ux = fabs(U(i, j, k));
          double vx = 0.25 * fabs(V(i, j, k) + V(i, j - 1, k) + V(i + 1, j, k) + V(i + 1, j - 1, k));
          double wx = 0.25 * fabs(W(i, j, k) + W(i, j, k - 1) + W(i + 1, j, k) + W(i + 1, j, k - 1));
          double dtix = ux / dxc[i] + vx / dyf[j] + wx / dzf[k];
          double uy = 0.25 * fabs(U(i, j, k) + U(i, j + 1, k) + U(i - 1, j + 1, k) + U(i - 1, j, k));
          double vy = fabs(V(i, j, k));
          double wy = 0.25 * fabs(W(i, j, k) + W(i, j + 1, k) + W(i, j + 1, k - 1) + W(i, j, k - 1));
          double dtiy = uy / dxf[i] + vy / dyc[j] + wy / dzf[k];
          double uz = 0.25 * fabs(U(i, j, k) + U(i - 1, j, k) + U(i - 1, j, k + 1) + U(i, j, k + 1));
          double vz = 0.25 * fabs(V(i, j, k) + V(i, j - 1, k) + V(i, j - 1, k + 1) + V(i, j, k + 1));
          double wz = fabs(W(i, j, k));
          double dtiz = uz / dxf[i] + vz / dyf[j] + wz / dzc[k];
#undef U
#undef V
#undef W
          if (dtix > dti) dti = dtix;
          if (dtiy > dti) dti = dtiy;
          if (dtiz > dti) dti = dtiz;
        }
    for (int d = 0; d < 3; d++)
      for (int q = 0; q <= B->n[d] + 1; q++)
        if (B->df[d][q] < dlmin) dlmin = B->df[d][q];
  }
  if (dti == 0.0) dti = 1.0;
  double dtmax = sqrt(3.0) / dti;
  if (c->impdiff) return dtmax;
  /* main.f90:660 passes max(visc, alpha_max) */
  double ve = c->nscal > 0 && c->alpha_max > c->visc ? c->alpha_max : c->visc;
  double dd = 1.65 / 12.0 / ve * dlmin * dlmin;
  return dtmax < dd ? dtmax : dd;
}

/* ------------------------------------------------------------------------ */
/* r2r transforms: FFTW kinds as documented (fft.f90:96-149 chooses them)    */
/* ------------------------------------------------------------------------ */
int oc_r2r_direct(int kind, int n, const double *x, double *y) {
  if (n < 1) return -1;
  for (int k = 0; k < n; k++) {
    double s = 0.0;
    switch (kind) {
    case OC_R2HC: {
      /* halfcomplex: r0..r_{n/2}, i_{(n+1)/2-1}..i_1 ; forward sign -1 */
      if (k <= n / 2) {
        for (int j = 0; j < n; j++) s += x[j] * cos(2.0 * PI * (double)j * k / n);
      } else {
        int kk = n - k;
        for (int j = 0; j < n; j++) s += -x[j] * sin(2.0 * PI * (double)j * kk / n);
      }
      break;
    }
    case OC_HC2R: {
      /* y_k = r0 + 2 sum_{j=1}^{(n-1)/2}(r_j cos - i_j sin) + r_{n/2}(-1)^k (n even) */
      s = x[0];
      for (int j = 1; j < (n + 1) / 2; j++)
        s += 2.0 * (x[j] * cos(2.0 * PI * (double)j * k / n) - x[n - j] * sin(2.0 * PI * (double)j * k / n));
      if (n % 2 == 0) s += x[n / 2] * ((k % 2) ? -1.0 : 1.0);
      break;
    }
    case OC_REDFT10:
      for (int j = 0; j < n; j++) s += 2.0 * x[j] * cos(PI * (j + 0.5) * k / n);
      break;
    case OC_REDFT01:
      s = x[0];
      for (int j = 1; j < n; j++) s += 2.0 * x[j] * cos(PI * j * (k + 0.5) / n);
      break;
    case OC_RODFT10:
      for (int j = 0; j < n; j++) s += 2.0 * x[j] * sin(PI * (j + 0.5) * (k + 1) / n);
      break;
    case OC_RODFT01:
      s = ((k % 2) ? -1.0 : 1.0) * x[n - 1];
      for (int j = 0; j < n - 1; j++) s += 2.0 * x[j] * sin(PI * (j + 1) * (k + 0.5) / n);
      break;
    case OC_REDFT11:
      for (int j = 0; j < n; j++) s += 2.0 * x[j] * cos(PI * (j + 0.5) * (k + 0.5) / n);
      break;
    case OC_RODFT11:
      for (int j = 0; j < n; j++) s += 2.0 * x[j] * sin(PI * (j + 0.5) * (k + 0.5) / n);
      break;
    case OC_REDFT00:
      if (n < 2) return -1;
      s = x[0] + ((k % 2) ? -1.0 : 1.0) * x[n - 1];
      for (int j = 1; j < n - 1; j++) s += 2.0 * x[j] * cos(PI * j * k / (n - 1));
      break;
    case OC_RODFT00:
      for (int j = 0; j < n; j++) s += 2.0 * x[j] * sin(PI * (j + 1) * (k + 1) / (n + 1));
      break;
    default: return -2;
    }
    y[k] = s;
  }
  return 0;
}

/* iterative radix-2 complex FFT, sign = -1 forward, +1 backward (unscaled)  */
static void cfft_pow2(int m, double *re, double *im, int sign) {
  for (int i = 1, j = 0; i < m; i++) {
    int bit = m >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) { double t = re[i]; re[i] = re[j]; re[j] = t; t = im[i]; im[i] = im[j]; im[j] = t; }
  }
  for (int len = 2; len <= m; len <<= 1) {
    double ang = sign * 2.0 * PI / len;
    for (int i = 0; i < m; i += len)
      for (int j = 0; j < len / 2; j++) {
        double wr = cos(ang * j), wi = sin(ang * j);
        int a = i + j, b = i + j + len / 2;
        double xr = re[b] * wr - im[b] * wi, xi = re[b] * wi + im[b] * wr;
        re[b] = re[a] - xr; im[b] = im[a] - xi;
        re[a] += xr; im[a] += xi;
      }
  }
}

static int is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

/* r2r through complex FFTs of the symmetric extension (n a power of two)   */
int oc_r2r(int kind, int n, const double *x, double *y) {
  if (!is_pow2(n) || n < 2 || kind == OC_REDFT00 || kind == OC_RODFT00) return oc_r2r_direct(kind, n, x, y);
  int m = 2 * n;
  double *re = xcalloc(m, sizeof(double)), *im = xcalloc(m, sizeof(double));
  switch (kind) {
  case OC_R2HC:
    for (int j = 0; j < n; j++) re[j] = x[j];
    cfft_pow2(n, re, im, -1);
    for (int k = 0; k <= n / 2; k++) y[k] = re[k];
    for (int k = 1; k < (n + 1) / 2; k++) y[n - k] = im[k];
    break;
  case OC_HC2R:
    re[0] = x[0];
    for (int k = 1; k < (n + 1) / 2; k++) {
      re[k] = x[k]; im[k] = x[n - k];
      re[n - k] = x[k]; im[n - k] = -x[n - k];
    }
    if (n % 2 == 0) re[n / 2] = x[n / 2];
    cfft_pow2(n, re, im, +1);
    for (int k = 0; k < n; k++) y[k] = re[k];
    break;
  case OC_REDFT10:
  case OC_RODFT10: {
    double sg = 1.0;
    for (int j = 0; j < n; j++) {
      double xv = (kind == OC_RODFT10) ? sg * x[j] : x[j];
      re[j] = xv; re[m - 1 - j] = xv;
      sg = -sg;
    }
    cfft_pow2(m, re, im, -1);
    for (int k = 0; k < n; k++) {
      double a = -PI * k / m; /* e^{-i pi k/(2n)} */
      double val = re[k] * cos(a) - im[k] * sin(a);
      if (kind == OC_REDFT10) y[k] = val; else y[n - 1 - k] = val;
    }
    break;
  }
  case OC_REDFT01:
  case OC_RODFT01: {
    for (int j = 0; j < n; j++) {
      double xv = (kind == OC_RODFT01) ? x[n - 1 - j] : x[j];
      double wgt = j == 0 ? 1.0 : 2.0;
      double a = PI * j / m;
      re[j] = wgt * xv * cos(a); im[j] = wgt * xv * sin(a);
    }
    cfft_pow2(m, re, im, +1);
    for (int k = 0; k < n; k++) y[k] = (kind == OC_RODFT01 && (k % 2)) ? -re[k] : re[k];
    break;
  }
  case OC_REDFT11:
  case OC_RODFT11: {
    for (int j = 0; j < n; j++) {
      double a = -PI * j / m;
      re[j] = x[j] * cos(a); im[j] = x[j] * sin(a);
    }
    cfft_pow2(m, re, im, -1);
    for (int k = 0; k < n; k++) {
      double a = -PI * (k + 0.5) / m;
      double zr = re[k] * cos(a) - im[k] * sin(a), zi = re[k] * sin(a) + im[k] * cos(a);
      y[k] = (kind == OC_REDFT11) ? 2.0 * zr : -2.0 * zi;
    }
    break;
  }
  default: free(re); free(im); return -2;
  }
  free(re); free(im);
  return 0;
}

/* find_fft fft.f90:96-149                                                   */
int oc_find_fft(const char *bc, int is_centered, int *kf, int *kb, double *norm) {
  char a = bc[0], b = bc[1];
  if (is_centered) {
    if (a == 'F' && b == 'F') { *kf = OC_R2HC; *kb = OC_HC2R; norm[0] = 1; norm[1] = 0; }
    else if (a == 'N' && b == 'N') { *kf = OC_REDFT10; *kb = OC_REDFT01; norm[0] = 2; norm[1] = 0; }
    else if (a == 'D' && b == 'D') { *kf = OC_RODFT10; *kb = OC_RODFT01; norm[0] = 2; norm[1] = 0; }
    else if (a == 'N' && b == 'D') { *kf = OC_REDFT11; *kb = OC_REDFT11; norm[0] = 2; norm[1] = 0; }
    else if (a == 'D' && b == 'N') { *kf = OC_RODFT11; *kb = OC_RODFT11; norm[0] = 2; norm[1] = 0; }
    else return -1;
  } else {
    if (a == 'F' && b == 'F') { *kf = OC_R2HC; *kb = OC_HC2R; norm[0] = 1; norm[1] = 0; }
    else if (a == 'N' && b == 'N') { *kf = OC_REDFT00; *kb = OC_REDFT00; norm[0] = 2; norm[1] = -1; }
    else if (a == 'D' && b == 'D') { *kf = OC_RODFT00; *kb = OC_RODFT00; norm[0] = 2; norm[1] = 1; }
    else if (a == 'N' && b == 'D') { *kf = OC_REDFT10; *kb = OC_REDFT01; norm[0] = 2; norm[1] = 0; }
    else if (a == 'D' && b == 'N') { *kf = OC_RODFT01; *kb = OC_RODFT10; norm[0] = 2; norm[1] = 0; }
    else return -1;
  }
  return 0;
}

/* eigenvalues fft.f90:151-189; the reference has no 'DN' case (its lambda
 * would stay undefined); DN has the spectrum of ND by reflection.            */
int oc_eigenvalues(int n, const char *bc, int is_centered, double *lam) {
  char a = bc[0], b = bc[1];
  if (a == 'F' && b == 'F') {
    for (int l = 1; l <= n; l++) { double s = sin((double)(l - 1) * PI / (double)n); lam[l - 1] = -4.0 * s * s; }
  } else if (a == 'N' && b == 'N') {
    for (int l = 1; l <= n; l++) {
      double s = is_centered ? sin((double)(l - 1) * PI / (2.0 * n)) : sin((double)(l - 1) * PI / (2.0 * (n - 1 + 1)));
      lam[l - 1] = -4.0 * s * s;
    }
  } else if (a == 'D' && b == 'D') {
    if (is_centered) {
      for (int l = 1; l <= n; l++) { double s = sin((double)l * PI / (2.0 * n)); lam[l - 1] = -4.0 * s * s; }
    } else {
      for (int l = 1; l <= n - 1; l++) { double s = sin((double)l * PI / (2.0 * (n + 1 - 1))); lam[l - 1] = -4.0 * s * s; }
      lam[n - 1] = 0.0;
    }
  } else if ((a == 'N' && b == 'D') || (a == 'D' && b == 'N')) {
    for (int l = 1; l <= n; l++) { double s = sin((double)(2 * l - 1) * PI / (4.0 * n)); lam[l - 1] = -4.0 * s * s; }
  } else return -1;
  return 0;
}

/* ------------------------------------------------------------------------ */
/* case set-up                                                               */
/* ------------------------------------------------------------------------ */
static void seterr(char *buf, int len, const char *msg, int blk) {
  if (buf && len > 0) snprintf(buf, (size_t)len, "%s (block %d)", msg, blk);
}

/* poiseuille_square initflow.f90:~355-367                                   */
static double poiseuille(double r, double lmin, double lmax) {
  double rr = (r - lmin) / (lmax - lmin);
  return 6.0 * (rr * (1.0 - rr));
}

static int assemble_op(oc_case *c, oop *o, char *err, int errlen);
static void free_op(oop *o);

oc_case *oc_create(int nblocks, const int *ng, const double *lmin, const double *lmax,
                   const int *gt, const double *gr, const char *cbcvel, const double *bcvel,
                   const char *cbcpre, const double *bcpre, const int *inflow_type,
                   int fft_axis, double visci, const double *bforce, char *errbuf, int errlen) {
  if (nblocks < 1 || nblocks > MAXB) { seterr(errbuf, errlen, "invalid number of blocks", nblocks); return NULL; }
  oc_case *c = xcalloc(1, sizeof(oc_case));
  c->nb = nblocks;
  c->fft_axis = fft_axis;
  c->visc = 1.0 / visci;
  for (int d = 0; d < 3; d++) c->bforce[d] = bforce[d];
  c->method = OC_SOLVE_AUTO; c->tol = 1e-13; c->maxiter = 20000;
  for (int ib = 0; ib < nblocks; ib++) {
    ob *B = &c->b[ib];
    for (int d = 0; d < 3; d++) {
      B->n[d] = ng[ib * 3 + d];
      B->lmin[d] = lmin[ib * 3 + d]; B->lmax[d] = lmax[ib * 3 + d];
      B->gt[d] = gt[ib * 3 + d]; B->gr[d] = gr[ib * 3 + d];
      if (B->n[d] < 1) { seterr(errbuf, errlen, "invalid grid size", ib + 1); oc_destroy(c); return NULL; }
      for (int s = 0; s < 2; s++) {
        B->cbcpre[d][s] = cbcpre[(ib * 3 + d) * 2 + s];
        B->bcpre[d][s] = bcpre[(ib * 3 + d) * 2 + s];
        B->inflow_type[d][s] = inflow_type[(ib * 3 + d) * 2 + s];
        for (int cp = 0; cp < 3; cp++) {
          B->cbcvel[cp][d][s] = cbcvel[((ib * 3 + cp) * 3 + d) * 2 + s];
          B->bcvel[cp][d][s] = bcvel[((ib * 3 + cp) * 3 + d) * 2 + s];
        }
      }
    }
  }
  /* is_periodic param.f90:220-225: FF pressure pairs on every block          */
  for (int d = 0; d < 3; d++) {
    c->is_periodic[d] = 1;
    for (int ib = 0; ib < nblocks; ib++)
      if (!(c->b[ib].cbcpre[d][0] == 'F' && c->b[ib].cbcpre[d][1] == 'F')) c->is_periodic[d] = 0;
  }
  /* block placement initmpi.f90:61-98 (BFS over 'F' links, block 1 at lo=1) */
  {
    int done[MAXB] = {0}, unl[MAXB] = {0};
    int ib = 0;
    c->b[0].lo_g[0] = c->b[0].lo_g[1] = c->b[0].lo_g[2] = 1;
    unl[0] = 1;
    for (;;) {
      if (ib < 0) { seterr(errbuf, errlen, "invalid connectivity", 0); oc_destroy(c); return NULL; }
      ob *B = &c->b[ib];
      for (int d = 0; d < 3; d++)
        for (int inb = 0; inb < 2; inb++) {
          if (B->cbcpre[d][inb] != 'F') continue;
          int fr = (int)lround(B->bcpre[d][inb]) - 1;
          if (fr < 0 || fr >= nblocks) { seterr(errbuf, errlen, "friend block does not exist", ib + 1); oc_destroy(c); return NULL; }
          if (unl[fr]) continue;
          if (inb == 0 && c->b[fr].lmin[d] > B->lmin[d]) continue;
          if (inb == 1 && c->b[fr].lmin[d] < B->lmin[d]) continue;
          int ioff = -(1 - inb) * c->b[fr].n[d] + inb * B->n[d];
          for (int e = 0; e < 3; e++) c->b[fr].lo_g[e] = B->lo_g[e];
          c->b[fr].lo_g[d] = B->lo_g[d] + ioff;
          unl[fr] = 1;
        }
      done[ib] = 1;
      int all = 1;
      for (int q = 0; q < nblocks; q++) all &= done[q];
      if (all) break;
      ib = -1;
      for (int q = 0; q < nblocks; q++) if (!done[q] && unl[q]) { ib = q; break; }
    }
  }
  /* periods initmpi.f90:104-113 */
  for (int d = 0; d < 3; d++) {
    int mn = 1 << 30, mx = -(1 << 30);
    for (int ib = 0; ib < nblocks; ib++) {
      if (c->b[ib].lo_g[d] < mn) mn = c->b[ib].lo_g[d];
      int hi = c->b[ib].lo_g[d] + c->b[ib].n[d] - 1;
      if (hi > mx) mx = hi;
    }
    c->periods[d] = c->is_periodic[d] ? mx - mn + 1 : 0;
  }
  /* neighbours initmpi.f90:150-239 (one piece per block) */
  for (int ib = 0; ib < nblocks; ib++) {
    ob *B = &c->b[ib];
    for (int d = 0; d < 3; d++)
      for (int s = 0; s < 2; s++) {
        B->nb[s][d] = -1;
        if (B->cbcpre[d][s] != 'F') continue;
        int fr = (int)lround(B->bcpre[d][s]) - 1;
        ob *F = &c->b[fr];
        int ok = F->cbcpre[d][1 - s] == 'F' && (int)lround(F->bcpre[d][1 - s]) - 1 == ib;
        for (int e = 0; e < 3; e++) {
          if (e == d) continue;
          ok = ok && F->lo_g[e] == B->lo_g[e] && F->n[e] == B->n[e] && F->lmin[e] == B->lmin[e] &&
               F->lmax[e] == B->lmax[e] && F->gr[e] == B->gr[e] && F->gt[e] == B->gt[e];
        }
        int blo = B->lo_g[d], bhi = B->lo_g[d] + B->n[d] - 1;
        int flo = F->lo_g[d], fhi = F->lo_g[d] + F->n[d] - 1;
        if (s == 0) ok = ok && (blo == fhi + 1 || c->periods[d] == fhi - blo + 1);
        else ok = ok && (bhi == flo - 1 || c->periods[d] == bhi - flo + 1);
        if (!ok) { seterr(errbuf, errlen, "expected connectivity between blocks is not possible", ib + 1); oc_destroy(c); return NULL; }
        B->nb[s][d] = fr;
      }
    for (int d = 0; d < 3; d++)
      for (int s = 0; s < 2; s++) B->is_bound[s][d] = B->nb[s][d] < 0;
  }
  /* vol_all param.f90:239-240 */
  c->vol_all = 0.0;
  for (int ib = 0; ib < nblocks; ib++) {
    ob *B = &c->b[ib];
    c->vol_all += (B->lmax[0] - B->lmin[0]) * (B->lmax[1] - B->lmin[1]) * (B->lmax[2] - B->lmin[2]);
  }
  /* grid initgrid.f90:9-83 then bound_grid 84-127 */
  for (int ib = 0; ib < nblocks; ib++) {
    ob *B = &c->b[ib];
    for (int d = 0; d < 3; d++) {
      int n = B->n[d];
      B->dc[d] = xcalloc(n + 2, sizeof(double)); B->df[d] = xcalloc(n + 2, sizeof(double));
      B->rc[d] = xcalloc(n + 2, sizeof(double)); B->rf[d] = xcalloc(n + 2, sizeof(double));
      oc_initgrid(n, B->gt[d], B->gr[d], B->lmin[d], B->lmax[d], B->dc[d], B->df[d], B->rc[d], B->rf[d]);
    }
  }
  {
    /* bound_grid: the ghost face spacing is the neighbour's first/last
     * interior spacing, all exchanged from the pre-exchange values          */
    double lov[MAXB][3], hiv[MAXB][3];
    for (int ib = 0; ib < nblocks; ib++)
      for (int d = 0; d < 3; d++) { lov[ib][d] = c->b[ib].df[d][1]; hiv[ib][d] = c->b[ib].df[d][c->b[ib].n[d]]; }
    for (int ib = 0; ib < nblocks; ib++) {
      ob *B = &c->b[ib];
      for (int d = 0; d < 3; d++) {
        int n = B->n[d];
        if (B->nb[0][d] >= 0) B->df[d][0] = hiv[B->nb[0][d]][d];
        if (B->nb[1][d] >= 0) B->df[d][n + 1] = lov[B->nb[1][d]][d];
        B->dc[d][0] = (B->df[d][1] + B->df[d][0]) / 2.0;
        B->dc[d][n] = (B->df[d][n + 1] + B->df[d][n]) / 2.0;
        B->dc[d][n + 1] = B->dc[d][n];
      }
    }
  }
  /* fields */
  for (int ib = 0; ib < nblocks; ib++) {
    ob *B = &c->b[ib];
    B->sx = B->n[0] + 2; B->sy = B->n[1] + 2;
    B->ntot = B->sx * B->sy * (B->n[2] + 2);
    for (int f = 0; f < 6; f++) B->f[f] = xcalloc(B->ntot, sizeof(double));
    long nin = (long)B->n[0] * B->n[1] * B->n[2];
    for (int f = 0; f < 3; f++) B->dto[f] = xcalloc(nin, sizeof(double));
  }
  /* inflow planes main.f90:600-636 with init_inflow initflow.f90:369-429 */
  for (int ib = 0; ib < nblocks; ib++) {
    ob *B = &c->b[ib];
    for (int d = 0; d < 3; d++)
      for (int s = 0; s < 2; s++) {
        B->is_inflow[d][s] = B->is_bound[s][d] && B->inflow_type[d][s] > 0 && B->cbcvel[d][d][s] == 'D';
        if (!B->is_inflow[d][s]) continue;
        int a1, a2; tang(d, &a1, &a2);
        int m1 = B->n[a1] + 2, m2 = B->n[a2] + 2;
        for (int cp = 0; cp < 3; cp++) B->vin[d][cp][s] = xcalloc((long)m1 * m2, sizeof(double));
        double uref = B->bcvel[d][d][s];
        int q1 = c->periods[a1] != 0 ? 0 : 1, q2 = c->periods[a2] != 0 ? 0 : 1;
        /* INFLOW_KOVASZNAY initflow.f90:407-416: reb = uref*lref/visc with
         * lref = 1 (main.f90:600-623), the profile taken at x0 = -0.5 */
        double pi = acos(-1.0), reb = uref * 1.0 / c->visc;
        double lambda = reb / 2 - sqrt(reb * reb / 4 + 4 * (pi * pi)), x0 = -0.5;
        for (int t2 = 0; t2 < m2; t2++)
          for (int t1 = 0; t1 < m1; t1++) {
            double val = 0.0;
            if (B->inflow_type[d][s] == 1) {
              double f1 = q1 ? poiseuille(B->rc[a1][t1], B->lmin[a1], B->lmax[a1]) : 1.0;
              double f2 = q2 ? poiseuille(B->rc[a2][t2], B->lmin[a2], B->lmax[a2]) : 1.0;
              val = uref * f1 * f2;
            } else if (B->inflow_type[d][s] == 2) {
              val = 1.0 - exp(lambda * x0) * cos(2 * pi * B->rc[a2][t2]);
              B->vin[d][a2][s][t1 + (long)m1 * t2] = lambda / (2.0 * pi) * exp(lambda * x0) * sin(2 * pi * B->rf[a2][t2]);
            }
            B->vin[d][d][s][t1 + (long)m1 * t2] = val;
          }
      }
  }
  /* BC rhs planes for the pressure: init_bc_rhs solver.f90:44-157           */
  for (int ib = 0; ib < nblocks; ib++) {
    ob *B = &c->b[ib];
    double dl[2][3];
    for (int d = 0; d < 3; d++) {
      /* dl = dxc_g(lo_g-1), dxc_g(hi_g) of the undistributed grid (main.f90:717-719) */
      /* dxc_g(lo_g-1) = drf_g(lo_g), dxc_g(hi_g) = drf_g(hi_g): the interior
       * face spacings, which bound_grid leaves untouched                   */
      dl[0][d] = B->df[d][1];
      dl[1][d] = B->df[d][B->n[d]];
    }
    for (int d = 0; d < 3; d++) {
      if (c->periods[d] != 0) continue;
      int a1, a2; tang(d, &a1, &a2);
      for (int s = 0; s < 2; s++) {
        if (!B->is_bound[s][d]) continue;
        double factor = 0.0;
        char ct = B->cbcpre[d][s];
        if (ct == 'N') { factor = dl[s][d] * B->bcpre[d][s]; if (s == 1) factor = -factor; }
        else if (ct == 'D') factor = -2.0 * B->bcpre[d][s];
        int q = s == 0 ? 1 : B->n[d];
        long m1 = B->n[a1], m2 = B->n[a2];
        B->rhsp[d][s] = xcalloc(m1 * m2, sizeof(double));
        for (int t2 = 1; t2 <= m2; t2++)
          for (int t1 = 1; t1 <= m1; t1++) {
            int ijk[3]; ijk[d] = q; ijk[a1] = t1; ijk[a2] = t2;
            double weight;
            switch (c->fft_axis) {
            case 1: weight = B->df[1][ijk[1]] * B->df[2][ijk[2]]; break;
            case 2: weight = B->df[0][ijk[0]] * B->df[2][ijk[2]]; break;
            case 3: weight = B->df[0][ijk[0]] * B->df[1][ijk[1]]; break;
            default: weight = B->df[0][ijk[0]] * B->df[1][ijk[1]] * B->df[2][ijk[2]]; break;
            }
            double rhs = factor * weight / (B->dc[d][q - (1 - s)] * B->df[d][q]);
            B->rhsp[d][s][(t1 - 1) + m1 * (t2 - 1)] = rhs;
          }
      }
    }
  }
  c->P.which = 0; c->P.stag = -1;
  if (assemble_op(c, &c->P, errbuf, errlen)) { oc_destroy(c); return NULL; }
  return c;
}

void oc_destroy(oc_case *c) {
  if (!c) return;
  for (int ib = 0; ib < c->nb; ib++) {
    ob *B = &c->b[ib];
    for (int d = 0; d < 3; d++) { free(B->dc[d]); free(B->df[d]); free(B->rc[d]); free(B->rf[d]); }
    for (int f = 0; f < 6; f++) free(B->f[f]);
    for (int f = 0; f < 3; f++) free(B->dto[f]);
    for (int is = 0; is < c->nscal; is++) { free(B->sv[is]); free(B->sdo[is]); }
    for (int d = 0; d < 3; d++)
      for (int s = 0; s < 2; s++) {
        free(B->rhsp[d][s]);
        for (int cp = 0; cp < 3; cp++) free(B->vin[d][cp][s]);
      }
  }
  free_op(&c->P);
  for (int q = 0; q < 3 + MAXS; q++) free_op(&c->H[q]);
  free(c);
}

int oc_nblocks(const oc_case *c) { return c->nb; }
double oc_vol_all(const oc_case *c) { return c->vol_all; }
void oc_block_info(const oc_case *c, int b, int *n3, int *lo3, int *nb6, int *per3) {
  const ob *B = &c->b[b];
  for (int d = 0; d < 3; d++) {
    n3[d] = B->n[d]; lo3[d] = B->lo_g[d]; per3[d] = c->periods[d];
    nb6[2 * d] = B->nb[0][d]; nb6[2 * d + 1] = B->nb[1][d];
  }
}
int oc_get_metric(const oc_case *c, int b, int which, double *out) {
  if (b < 0 || b >= c->nb || which < 0 || which > 11) return -1;
  const ob *B = &c->b[b];
  int d = (which % 6) / 2, cf = which % 2; /* 0 = c, 1 = f */
  const double *src = which < 6 ? (cf == 0 ? B->dc[d] : B->df[d]) : (cf == 0 ? B->rc[d] : B->rf[d]);
  memcpy(out, src, sizeof(double) * (B->n[d] + 2));
  return 0;
}
/* field id -> (array, ghosted?) including scalars 9+2i / dsdtrko 10+2i     */
static double *field_of(const oc_case *c, int b, int fid, int *ghosted) {
  if (b < 0 || b >= c->nb || fid < 0) return NULL;
  ob *B = (ob *)&c->b[b];
  if (fid < 6) { *ghosted = 1; return B->f[fid]; }
  if (fid < 9) { *ghosted = 0; return B->dto[fid - 6]; }
  int is = (fid - 9) / 2;
  if (is >= c->nscal) return NULL;
  *ghosted = (fid - 9) % 2 == 0;
  return *ghosted ? B->sv[is] : B->sdo[is];
}
int oc_field_size(const oc_case *c, int b, int fid) {
  int gh = 0;
  if (!field_of(c, b, fid, &gh)) return -1;
  const ob *B = &c->b[b];
  return gh ? (int)B->ntot : B->n[0] * B->n[1] * B->n[2];
}
int oc_get_field(const oc_case *c, int b, int fid, double *out) {
  int n = oc_field_size(c, b, fid), gh;
  if (n < 0) return -1;
  memcpy(out, field_of(c, b, fid, &gh), sizeof(double) * n);
  return 0;
}
int oc_set_field(oc_case *c, int b, int fid, const double *in) {
  int n = oc_field_size(c, b, fid), gh;
  if (n < 0) return -1;
  memcpy(field_of(c, b, fid, &gh), in, sizeof(double) * n);
  return 0;
}

/* scalar inputs: &scalar of dns.nml (param.f90:96-100,148-163) and
 * block_cbcscal/block_bcscal of blocks.nml (flat [b][iscal][dir][side])    */
int oc_set_buoyancy(oc_case *c, int enable) {
  if (enable && c->nscal <= 0) return -1;
  c->buoyancy = enable != 0;
  return 0;
}

int oc_set_scalars(oc_case *c, int nscal, const double *alphai, double beta, const double *gacc,
                   const double *ssource, const int *is_sforced, const double *scalf,
                   const char *cbcscal, const double *bcscal) {
  if (nscal < 0 || nscal > MAXS) return -1;
  for (int ib = 0; ib < c->nb; ib++)
    for (int is = 0; is < c->nscal; is++) { free(c->b[ib].sv[is]); free(c->b[ib].sdo[is]); }
  c->nscal = nscal;
  c->beta = beta;
  for (int q = 0; q < 3; q++) c->gacc[q] = gacc[q];
  c->alpha_max = 0.0;
  for (int is = 0; is < nscal; is++) {
    c->salpha[is] = 1.0 / alphai[is];
    if (is == 0 || c->salpha[is] > c->alpha_max) c->alpha_max = c->salpha[is];
    c->ssource[is] = ssource[is];
    c->is_sforced[is] = is_sforced[is];
    c->scalf[is] = scalf[is];
    c->sf[is] = 0.0;
  }
  for (int ib = 0; ib < c->nb; ib++) {
    ob *B = &c->b[ib];
    long nin = (long)B->n[0] * B->n[1] * B->n[2];
    for (int is = 0; is < nscal; is++) {
      B->sv[is] = xcalloc(B->ntot, sizeof(double));
      B->sdo[is] = xcalloc(nin, sizeof(double));
      for (int d = 0; d < 3; d++)
        for (int sd = 0; sd < 2; sd++) {
          long q = (((long)ib * nscal + is) * 3 + d) * 2 + sd;
          B->cbcsca[is][d][sd] = cbcscal[q];
          B->bcsca[is][d][sd] = bcscal[q];
        }
    }
  }
  return 0;
}

/* rk_scal rk.f90:130-204 with scal_d / scal_a scal.f90:89-155, then
 * bulk_mean_s / bulk_forcing_s scal.f90:157-199 (main.f90:908-913)          */
void oc_rk_scal(oc_case *c, int irk, double dt, double *fout) {
  const double f1 = RKC[irk - 1][0] * dt, f2 = RKC[irk - 1][1] * dt, f12 = f1 + f2;
  for (int is = 0; is < c->nscal; is++) {
    const double alpha = c->salpha[is];
    /* _IMPDIFF: the diffusion goes to dsdtrkd (rk.f90:151-158) */
    double *dsd[MAXB] = {NULL};
    for (int ib = 0; ib < c->nb; ib++) {
      ob *B = &c->b[ib];
      const int nx = B->n[0], ny = B->n[1], nz = B->n[2];
      const double *u = B->f[0], *v = B->f[1], *w = B->f[2];
      double *s = B->sv[is], *dso = B->sdo[is];
      const double *dxc = B->dc[0], *dxf = B->df[0], *dyc = B->dc[1], *dyf = B->df[1],
                   *dzc = B->dc[2], *dzf = B->df[2];
      double *ds = xcalloc((long)nx * ny * nz, sizeof(double));
      double *dd = c->impdiff ? (dsd[ib] = xcalloc((long)nx * ny * nz, sizeof(double))) : NULL;
#define S(i, j, k) s[IDX(B, i, j, k)]
#define U(i, j, k) u[IDX(B, i, j, k)]
#define V(i, j, k) v[IDX(B, i, j, k)]
#define W(i, j, k) w[IDX(B, i, j, k)]
#pragma omp parallel for collapse(2) schedule(static)
      for (int k = 1; k <= nz; k++)
        for (int j = 1; j <= ny; j++)
          for (int i = 1; i <= nx; i++) {
            double d = 0.0;
            double dsdxp = (S(i + 1, j, k) - S(i, j, k)) / dxc[i];
            double dsdxm = (S(i, j, k) - S(i - 1, j, k)) / dxc[i - 1];
            double dsdyp = (S(i, j + 1, k) - S(i, j, k)) / dyc[j];
            double dsdym = (S(i, j, k) - S(i, j - 1, k)) / dyc[j - 1];
            double dsdzp = (S(i, j, k + 1) - S(i, j, k)) / dzc[k];
            double dsdzm = (S(i, j, k) - S(i, j, k - 1)) / dzc[k - 1];
            d = d + alpha * (dsdxp - dsdxm) / dxf[i] + alpha * (dsdyp - dsdym) / dyf[j] +
                alpha * (dsdzp - dsdzm) / dzf[k];
            if (dd) { dd[IIDX(B, i, j, k)] = d; d = 0.0; }
            double usim = 0.5 * (S(i - 1, j, k) + S(i, j, k)) * U(i - 1, j, k);
            double usip = 0.5 * (S(i + 1, j, k) + S(i, j, k)) * U(i, j, k);
            double vsjm = 0.5 * (S(i, j - 1, k) + S(i, j, k)) * V(i, j - 1, k);
            double vsjp = 0.5 * (S(i, j + 1, k) + S(i, j, k)) * V(i, j, k);
            double wskm = 0.5 * (S(i, j, k - 1) + S(i, j, k)) * W(i, j, k - 1);
            double wskp = 0.5 * (S(i, j, k + 1) + S(i, j, k)) * W(i, j, k);
            d = d + (-usip + usim) / dxf[i] + (-vsjp + vsjm) / dyf[j] + (-wskp + wskm) / dzf[k];
            ds[IIDX(B, i, j, k)] = d;
          }
#pragma omp parallel for collapse(2) schedule(static)
      for (int k = 1; k <= nz; k++)
        for (int j = 1; j <= ny; j++)
          for (int i = 1; i <= nx; i++) {
            long q = IIDX(B, i, j, k);
            S(i, j, k) = S(i, j, k) + f1 * ds[q] + f2 * dso[q] + f12 * c->ssource[is];
            if (dd) S(i, j, k) = S(i, j, k) + f12 * dd[q];
            dso[q] = ds[q];
          }
#undef S
#undef U
#undef V
#undef W
      free(ds);
    }
    double f = 0.0;
    if (c->is_sforced[is]) {
      double mean = 0.0;
      for (int ib = 0; ib < c->nb; ib++) {
        ob *B = &c->b[ib];
        for (int k = 1; k <= B->n[2]; k++)
          for (int j = 1; j <= B->n[1]; j++)
            for (int i = 1; i <= B->n[0]; i++)
              mean += B->sv[is][IDX(B, i, j, k)] * B->df[0][i] * B->df[1][j] * B->df[2][k];
      }
      mean = mean / c->vol_all;
      f = c->scalf[is] - mean;
    }
    if (c->impdiff) /* rhs of the Helmholtz equation rk.f90:190-202, after the mean */
      for (int ib = 0; ib < c->nb; ib++) {
        ob *B = &c->b[ib];
        for (int k = 1; k <= B->n[2]; k++)
          for (int j = 1; j <= B->n[1]; j++)
            for (int i = 1; i <= B->n[0]; i++)
              B->sv[is][IDX(B, i, j, k)] = B->sv[is][IDX(B, i, j, k)] - .5 * f12 * dsd[ib][IIDX(B, i, j, k)];
        free(dsd[ib]);
      }
    if (c->is_sforced[is]) {
      for (int ib = 0; ib < c->nb; ib++) {
        ob *B = &c->b[ib];
        for (int k = 1; k <= B->n[2]; k++)
          for (int j = 1; j <= B->n[1]; j++)
            for (int i = 1; i <= B->n[0]; i++) B->sv[is][IDX(B, i, j, k)] += f;
      }
    }
    c->sf[is] = f;
    if (fout) fout[is] = f;
  }
}

/* ------------------------------------------------------------------------ */
/* Elliptic operators: the pressure Poisson operator and, with implicit     */
/* diffusion, the Helmholtz operator of each velocity component and scalar. */
/* init_matrix_2d solver.f90:687-842 + init_n_2d_matrices 981-1020 (lambda  */
/* on the diagonal), init_matrix_3d solver.f90:158-372, init_bc_rhs 44-157. */
/* A component is staggered along its own axis (is_centered = .false.):     */
/* its unknowns sit on the faces lo..hi (lo..hi-1 when the block is bounded */
/* at hi, main.f90:774-775), couplings use dx1 = the face spacing dxf       */
/* shifted by qqq = 1, weights use dx2 = dxc, and a Dirichlet value lies on */
/* the boundary face itself (sgn = 0, factor = -bc).                        */
/* ------------------------------------------------------------------------ */
static char op_cbc(const oc_case *c, const oop *o, const ob *B, int d, int s) {
  (void)c;
  if (o->which == 0) return B->cbcpre[d][s];
  if (o->which <= 3) return B->cbcvel[o->which - 1][d][s];
  return B->cbcsca[o->which - 4][d][s];
}
static double op_bcval(const oop *o, const ob *B, int d, int s) {
  if (o->which == 0) return B->bcpre[d][s];
  if (o->which <= 3) return B->bcvel[o->which - 1][d][s];
  return B->bcsca[o->which - 4][d][s];
}
/* spacing to the lower / upper neighbour of unknown i (1-based) along a,   */
/* and the cell width dx2 along a                                           */
static inline double sp_lo(const oop *o, const ob *B, int a, int i) { return o->stag == a ? B->df[a][i] : B->dc[a][i - 1]; }
static inline double sp_hi(const oop *o, const ob *B, int a, int i) { return o->stag == a ? B->df[a][i + 1] : B->dc[a][i]; }
static inline double wdx2(const oop *o, const ob *B, int a, int i) { return o->stag == a ? B->dc[a][i] : B->df[a][i]; }
/* the operator's weight of one interior cell: dx2 over the non-FFT axes   */
static double op_weight(const oc_case *c, const oop *o, const ob *B, const int ijk[3]) {
  double w = 1.0;
  for (int a = 0; a < 3; a++)
    if (a != c->fft_axis - 1) w *= wdx2(o, B, a, ijk[a]);
  return w;
}

static int assemble_op(oc_case *c, oop *o, char *err, int errlen) {
  int fa = c->fft_axis;
  for (int ib = 0; ib < c->nb; ib++) {
    ob *B = &c->b[ib];
    for (int a = 0; a < 3; a++)
      o->ext[ib][a] = B->n[a] - ((o->stag == a && B->is_bound[1][a]) ? 1 : 0);
  }
  if (fa >= 1 && fa <= 3) {
    int d = fa - 1;
    o->nd = 2;
    int k = 0;
    for (int e = 0; e < 3; e++) if (e != d) o->ax[k++] = e;
    ob *B0 = &c->b[0];
    char bc2[2] = {op_cbc(c, o, B0, d, 0), op_cbc(c, o, B0, d, 1)};
    int centred = o->stag != d;
    /* fftini fft.f90:31-65: a staggered non-periodic FFT axis transforms its
     * first n-1 points.  eigenvalues fft.f90:151-189 does not match those
     * transforms for the staggered NN, ND and DN pairs (its formulas are the
     * n-point ones, DN has none): only FF and DD are taken on a staggered
     * FFT axis.                                                              */
    o->qfft = (!centred && !(bc2[0] == 'F' && bc2[1] == 'F')) ? 1 : 0;
    if (!centred && o->qfft && !(bc2[0] == 'D' && bc2[1] == 'D')) {
      if (err) snprintf(err, (size_t)errlen, "staggered %c%c pair along the FFT axis is not supported", bc2[0], bc2[1]);
      return -1;
    }
    double norm[2];
    if (oc_find_fft(bc2, centred, &o->kind_fwd, &o->kind_bwd, norm)) {
      if (err) snprintf(err, (size_t)errlen, "unsupported boundary pair %c%c along the FFT axis", bc2[0], bc2[1]);
      return -1;
    }
    int n = B0->n[d];
    o->nmodes = n - o->qfft;
    o->normfft = 1.0 / (norm[0] * (n + norm[1] - o->qfft));
    o->lambda = xcalloc(n, sizeof(double));
    oc_eigenvalues(n, bc2, centred, o->lambda);
    /* dl(0,idir) of main.f90:717-728 / 769-771: dxc_g(lo_g-1) = dxf_g(lo_g)
     * for the pressure, dxf_g(lo_g) for a staggered axis                    */
    double dl = B0->df[d][1];
    for (int m = 0; m < n; m++) o->lambda[m] /= dl * dl;
  } else {
    o->nd = 3; o->ax[0] = 0; o->ax[1] = 1; o->ax[2] = 2;
    o->nmodes = 1; o->qfft = 0;
    o->lambda = xcalloc(1, sizeof(double));
    o->normfft = 1.0;
  }
  int nd = o->nd;
  long nunk = 0;
  for (int ib = 0; ib < c->nb; ib++) {
    o->off[ib] = nunk;
    long cnt = 1;
    for (int e = 0; e < nd; e++) cnt *= o->ext[ib][o->ax[e]];
    nunk += cnt;
  }
  o->nunk = nunk;
  o->diag0 = xcalloc(nunk, sizeof(double));
  o->wgt = xcalloc(nunk, sizeof(double));
  o->coef = xcalloc(nunk * 2 * nd, sizeof(double));
  o->nbr = xcalloc(nunk * 2 * nd, sizeof(long));
  o->has_dirichlet = 0;
  for (int ib = 0; ib < c->nb; ib++) {
    ob *B = &c->b[ib];
    int nn[3] = {1, 1, 1};
    for (int e = 0; e < nd; e++) nn[e] = o->ext[ib][o->ax[e]];
    for (int i3 = 1; i3 <= nn[2]; i3++)
      for (int i2 = 1; i2 <= nn[1]; i2++)
        for (int i1 = 1; i1 <= nn[0]; i1++) {
          int idx[3] = {i1, i2, i3};
          long q = o->off[ib] + (i1 - 1) + (long)nn[0] * ((i2 - 1) + (long)nn[1] * (i3 - 1));
          double weight = 1.0;
          for (int e = 0; e < nd; e++) weight *= wdx2(o, B, o->ax[e], idx[e]);
          double cm[3], cp[3];
          for (int e = 0; e < nd; e++) {
            double other = 1.0;
            for (int oo = 0; oo < nd; oo++) if (oo != e) other *= wdx2(o, B, o->ax[oo], idx[oo]);
            cm[e] = 1.0 / sp_lo(o, B, o->ax[e], idx[e]) * other;
            cp[e] = 1.0 / sp_hi(o, B, o->ax[e], idx[e]) * other;
          }
          double ssum = 0.0;
          for (int e = 0; e < nd; e++) ssum += cm[e] + cp[e];
          double cc = -ssum;
          for (int e = 0; e < nd; e++) {
            int a = o->ax[e];
            if (c->periods[a] != 0) continue;
            for (int s = 0; s < 2; s++) {
              if (!B->is_bound[s][a] || idx[e] != (s == 0 ? 1 : nn[e])) continue;
              char ct = op_cbc(c, o, B, a, s);
              double sgn = ct == 'N' ? 1.0 : (ct == 'D' ? (o->stag == a ? 0.0 : -1.0) : 0.0);
              if (ct == 'D') o->has_dirichlet = 1;
              if (s == 0) { cc = cc + sgn * cm[e]; cm[e] = 0.0; }
              else { cc = cc + sgn * cp[e]; cp[e] = 0.0; }
            }
          }
          o->diag0[q] = cc;
          o->wgt[q] = weight;
          for (int e = 0; e < nd; e++) {
            int a = o->ax[e];
            for (int s = 0; s < 2; s++) {
              long slot = q * 2 * nd + 2 * e + s;
              double cf = s == 0 ? cm[e] : cp[e];
              o->coef[slot] = cf;
              o->nbr[slot] = -1;
              if (cf == 0.0) continue;
              int j2[3] = {idx[0], idx[1], idx[2]};
              int tb = ib;
              if (s == 0 && idx[e] == 1) { tb = B->nb[0][a]; j2[e] = o->ext[tb][a]; }
              else if (s == 1 && idx[e] == nn[e]) { tb = B->nb[1][a]; j2[e] = 1; }
              else j2[e] += s == 0 ? -1 : 1;
              int tn[3] = {1, 1, 1};
              for (int oo = 0; oo < nd; oo++) tn[oo] = o->ext[tb][o->ax[oo]];
              /* the struct grid of solver.f90:221-224 is the union of the
               * blocks' extents: a link into a neighbour's excluded bounded
               * staggered face (e.g. v at the top of a lower block, whose
               * x-neighbour is the wall face of the block beside it) leaves
               * the grid and meets a zero value, so it drops out            */
              int outside = 0;
              for (int oo = 0; oo < nd; oo++) if (j2[oo] > tn[oo]) outside = 1;
              if (outside) { o->coef[slot] = 0.0; continue; }
              o->nbr[slot] = o->off[tb] + (j2[0] - 1) + (long)tn[0] * ((j2[1] - 1) + (long)tn[1] * (j2[2] - 1));
            }
          }
        }
  }
  o->work = xcalloc(nunk * 6, sizeof(double));
  /* init_bc_rhs solver.f90:44-157 for the Helmholtz operators (the pressure
   * keeps its planes in ob.rhsp): over the unknowns of each bounded face,
   * factor * weight / (dx1(i-(1-ib)+qqq) * dx2(i)), times bc_array / bc
   * where an inflow array gives the boundary values (main.f90:600-636)     */
  if (o->which > 0) {
    for (int ib = 0; ib < c->nb; ib++) {
      ob *B = &c->b[ib];
      for (int d = 0; d < 3; d++) {
        if (c->periods[d] != 0) continue;
        int a1, a2; tang(d, &a1, &a2);
        for (int s = 0; s < 2; s++) {
          if (!B->is_bound[s][d]) continue;
          char ct = op_cbc(c, o, B, d, s);
          double bcv = op_bcval(o, B, d, s), factor = 0.0;
          /* dl: the block's first / last spacing (dxc_g(lo_g-1), dxc_g(hi_g);
           * dxf_g(lo_g), dxf_g(hi_g) along the staggered axis)              */
          double dl = s == 0 ? B->df[d][1] : B->df[d][B->n[d]];
          if (ct == 'N') { factor = dl * bcv; if (s == 1) factor = -factor; }
          else if (ct == 'D') factor = o->stag == d ? -bcv : -2.0 * bcv;
          int q = s == 0 ? 1 : o->ext[ib][d];
          long m1 = o->ext[ib][a1], m2 = o->ext[ib][a2];
          const double *arr = NULL;
          if (o->which <= 3 && B->is_inflow[d][s]) arr = B->vin[d][o->which - 1][s];
          o->rhs[ib][d][s] = xcalloc(m1 * m2, sizeof(double));
          for (int t2 = 1; t2 <= m2; t2++)
            for (int t1 = 1; t1 <= m1; t1++) {
              int ijk[3]; ijk[d] = q; ijk[a1] = t1; ijk[a2] = t2;
              double spc = s == 0 ? sp_lo(o, B, d, q) : sp_hi(o, B, d, q);
              double rhs = factor * op_weight(c, o, B, ijk) / (spc * wdx2(o, B, d, q));
              if (arr && bcv != 0.0) rhs = rhs * arr[t1 + (long)(B->n[a1] + 2) * t2] / bcv;
              o->rhs[ib][d][s][(t1 - 1) + m1 * (t2 - 1)] = rhs;
            }
        }
      }
    }
  }
  return 0;
}

static void free_op(oop *o) {
  free(o->lambda); free(o->diag0); free(o->wgt); free(o->coef); free(o->nbr);
  free(o->perm); free(o->iperm);
  for (int s = 0; s < 3; s++)
    if (o->bandL[s]) { for (int m = 0; m < o->nmodes; m++) free(o->bandL[s][m]); free(o->bandL[s]); }
  free(o->t_lambda[0]); free(o->t_lambda[1]); free(o->work);
  for (int ib = 0; ib < MAXB; ib++)
    for (int d = 0; d < 3; d++) { free(o->rhs[ib][d][0]); free(o->rhs[ib][d][1]); }
  memset(o, 0, sizeof(*o));
}

int oc_num_unknowns(const oc_case *c) { return (int)c->P.nunk; }
int oc_get_lambda(const oc_case *c, double *lambda, double *normfft) {
  if (lambda) memcpy(lambda, c->P.lambda, sizeof(double) * c->P.nmodes);
  if (normfft) *normfft = c->P.normfft;
  return c->P.nmodes;
}

static inline double op_lam(const oop *o, int mode) { return o->lambda[mode] + o->shift; }

static void apply_op(const oop *o, int mode, const double *x, double *y) {
  int nd = o->nd;
  double lam = op_lam(o, mode);
#pragma omp parallel for schedule(static)
  for (long q = 0; q < o->nunk; q++) {
    double s = (o->diag0[q] + lam * o->wgt[q]) * x[q];
    for (int t = 0; t < 2 * nd; t++) {
      long nb = o->nbr[q * 2 * nd + t];
      if (nb >= 0) s += o->coef[q * 2 * nd + t] * x[nb];
    }
    y[q] = s;
  }
}

int oc_apply_operator(const oc_case *c, int mode, const double *x, double *y) {
  if (mode < 0 || mode >= c->P.nmodes) return -1;
  apply_op(&c->P, mode, x, y);
  return 0;
}

static int mode_singular(const oop *o, int mode) {
  return !o->has_dirichlet && op_lam(o, mode) == 0.0;
}

/* Jacobi-preconditioned CG on -A (A is negative semi-definite)             */
static int solve_cg(const oc_case *c, oop *o, int mode, const double *b_in, double *x) {
  long n = o->nunk;
  double *r = o->work, *z = r + n, *p = z + n, *q = p + n, *b = q + n;
  double lam = op_lam(o, mode);
  memcpy(b, b_in, sizeof(double) * n);
  if (mode_singular(o, mode)) {
    double m = 0; for (long i = 0; i < n; i++) m += b[i]; m /= n;
    for (long i = 0; i < n; i++) b[i] -= m;
  }
  double bn = 0; for (long i = 0; i < n; i++) bn += b[i] * b[i];
  bn = sqrt(bn);
  if (bn == 0.0) { memset(x, 0, sizeof(double) * n); return 0; }
  apply_op(o, mode, x, q);
  for (long i = 0; i < n; i++) r[i] = -(b[i] - q[i]); /* residual of -A x = -b */
  double rz = 0;
  for (long i = 0; i < n; i++) { z[i] = r[i] / (-(o->diag0[i] + lam * o->wgt[i])); p[i] = z[i]; rz += r[i] * z[i]; }
  int it;
  for (it = 0; it < c->maxiter; it++) {
    double rn = 0; for (long i = 0; i < n; i++) rn += r[i] * r[i];
    if (sqrt(rn) <= c->tol * bn) break;
    apply_op(o, mode, p, q);
    double pq = 0; for (long i = 0; i < n; i++) { q[i] = -q[i]; pq += p[i] * q[i]; }
    double al = rz / pq;
    for (long i = 0; i < n; i++) { x[i] += al * p[i]; r[i] -= al * q[i]; }
    double rz2 = 0;
    for (long i = 0; i < n; i++) { z[i] = r[i] / (-(o->diag0[i] + lam * o->wgt[i])); rz2 += r[i] * z[i]; }
    double be = rz2 / rz; rz = rz2;
    for (long i = 0; i < n; i++) p[i] = z[i] + be * p[i];
  }
  return it;
}

/* banded Cholesky of -A in an ordering of the union by global coordinates  */
static int band_setup(const oc_case *c, oop *o) {
  if (o->perm) return o->band_ok;
  long n = o->nunk; int nd = o->nd;
  o->perm = xcalloc(n, sizeof(long)); o->iperm = xcalloc(n, sizeof(long));
  /* global coords of each unknown; order: highest reduced axis that is not
   * periodic is the major key (wrap couplings stay inside the band)         */
  long *key = xcalloc(n, sizeof(long));
  int order[3] = {0, 1, 2};
  if (nd == 2) { order[0] = 0; order[1] = 1; if (c->periods[o->ax[1]] != 0 && c->periods[o->ax[0]] == 0) { order[0] = 1; order[1] = 0; } }
  else {
    int maj = 2;
    for (int e = 2; e >= 0; e--) if (c->periods[o->ax[e]] == 0) { maj = e; break; }
    int k = 0; for (int e = 0; e < 3; e++) if (e != maj) order[k++] = e;
    order[2] = maj;
  }
  int gmin[3] = {1 << 30, 1 << 30, 1 << 30}, gmax[3] = {-(1 << 30), -(1 << 30), -(1 << 30)};
  for (int ib = 0; ib < c->nb; ib++)
    for (int e = 0; e < nd; e++) {
      int a = o->ax[e];
      if (c->b[ib].lo_g[a] < gmin[e]) gmin[e] = c->b[ib].lo_g[a];
      int hi = c->b[ib].lo_g[a] + o->ext[ib][a] - 1;
      if (hi > gmax[e]) gmax[e] = hi;
    }
  long ext[3] = {1, 1, 1};
  for (int e = 0; e < nd; e++) ext[e] = gmax[e] - gmin[e] + 1;
  for (int ib = 0; ib < c->nb; ib++) {
    const ob *B = &c->b[ib];
    int nn[3] = {1, 1, 1};
    for (int e = 0; e < nd; e++) nn[e] = o->ext[ib][o->ax[e]];
    for (int i3 = 0; i3 < nn[2]; i3++)
      for (int i2 = 0; i2 < nn[1]; i2++)
        for (int i1 = 0; i1 < nn[0]; i1++) {
          long q = o->off[ib] + i1 + (long)nn[0] * (i2 + (long)nn[1] * i3);
          int loc[3] = {i1, i2, i3};
          long g[3] = {0, 0, 0};
          for (int e = 0; e < nd; e++) g[e] = B->lo_g[o->ax[e]] + loc[e] - gmin[e];
          long kk = 0;
          for (int oo = nd - 1; oo >= 0; oo--) kk = kk * ext[order[oo]] + g[order[oo]];
          key[q] = kk;
        }
  }
  long kmax = 1; for (int e = 0; e < nd; e++) kmax *= ext[e];
  long *slot = xcalloc(kmax, sizeof(long));
  for (long k = 0; k < kmax; k++) slot[k] = -1;
  for (long q = 0; q < n; q++) slot[key[q]] = q;
  long pos = 0;
  for (long k = 0; k < kmax; k++) if (slot[k] >= 0) { o->perm[pos] = slot[k]; o->iperm[slot[k]] = pos; pos++; }
  free(slot); free(key);
  long bw = 0;
  for (long q = 0; q < n; q++)
    for (int t = 0; t < 2 * nd; t++) {
      long nb = o->nbr[q * 2 * nd + t];
      if (nb < 0) continue;
      long dd = labs(o->iperm[q] - o->iperm[nb]);
      if (dd > bw) bw = dd;
    }
  o->bw = bw;
  o->band_ok = (double)n * (double)(bw + 1) * (double)o->nmodes < 4.0e8;
  return o->band_ok;
}

/* factorisations are kept per shift (the three RK substages of a Helmholtz
 * operator have three shifts; the pressure has one)                        */
static double **band_slot(oop *o) {
  for (int s = 0; s < 3; s++)
    if (o->bandL[s] && o->band_shift[s] == o->shift) return o->bandL[s];
  int s = o->band_next;
  o->band_next = (o->band_next + 1) % 3;
  if (o->bandL[s]) { for (int m = 0; m < o->nmodes; m++) free(o->bandL[s][m]); free(o->bandL[s]); }
  o->bandL[s] = xcalloc(o->nmodes, sizeof(double *));
  o->band_shift[s] = o->shift;
  return o->bandL[s];
}

static int band_factor(oop *o, double **slot, int mode) {
  long n = o->nunk, bw = o->bw; int nd = o->nd;
  double *L = xcalloc(n * (bw + 1), sizeof(double)); /* L[i*(bw+1) + (i-j)] */
  double lam = op_lam(o, mode);
  int sing = mode_singular(o, mode);
  for (long q = 0; q < n; q++) {
    long i = o->iperm[q];
    L[i * (bw + 1)] += -(o->diag0[q] + lam * o->wgt[q]);
    for (int t = 0; t < 2 * nd; t++) {
      long nb = o->nbr[q * 2 * nd + t];
      if (nb < 0) continue;
      long j = o->iperm[nb];
      if (j < i) L[i * (bw + 1) + (i - j)] += -o->coef[q * 2 * nd + t];
    }
  }
  if (sing) { /* pin the first unknown of the ordering */
    for (long t = 1; t <= bw && t < n; t++) L[t * (bw + 1) + t] = 0.0;
    L[0] = 1.0;
  }
  for (long i = 0; i < n; i++) {
    long j0 = i - bw < 0 ? 0 : i - bw;
    for (long j = j0; j <= i; j++) {
      double s = L[i * (bw + 1) + (i - j)];
      long k0 = j0 > j - bw ? j0 : j - bw;
      if (k0 < 0) k0 = 0;
      for (long k = k0; k < j; k++) s -= L[i * (bw + 1) + (i - k)] * L[j * (bw + 1) + (j - k)];
      if (i == j) {
        if (s <= 0.0) { free(L); return -1; }
        L[i * (bw + 1)] = sqrt(s);
      } else L[i * (bw + 1) + (i - j)] = s / L[j * (bw + 1)];
    }
  }
  slot[mode] = L;
  return 0;
}

static int solve_band(oop *o, int mode, const double *b, double *x) {
  double **slot = band_slot(o);
  if (!slot[mode] && band_factor(o, slot, mode)) return -1;
  long n = o->nunk, bw = o->bw;
  const double *L = slot[mode];
  double *y = o->work;
  for (long i = 0; i < n; i++) y[i] = -b[o->perm[i]];
  if (mode_singular(o, mode)) y[0] = 0.0;
  for (long i = 0; i < n; i++) {
    double s = y[i];
    long k0 = i - bw < 0 ? 0 : i - bw;
    for (long k = k0; k < i; k++) s -= L[i * (bw + 1) + (i - k)] * y[k];
    y[i] = s / L[i * (bw + 1)];
  }
  for (long i = n - 1; i >= 0; i--) {
    double s = y[i];
    long k1 = i + bw >= n ? n - 1 : i + bw;
    for (long k = i + 1; k <= k1; k++) s -= L[k * (bw + 1) + (k - i)] * y[k];
    y[i] = s / L[i * (bw + 1)];
  }
  for (long i = 0; i < n; i++) x[o->perm[i]] = y[i];
  return 0;
}

/* tensor solver: r2r along one uniform in-plane axis (two for 3D) and a
 * tridiagonal solve along the remaining axis; single-block cases and the
 * pressure operator only in the oracle.                                     */
static int tensor_setup(oc_case *c, oop *o) {
  if (o->tensor_ok) return o->tensor_ok > 0;
  o->tensor_ok = -1;
  if (c->nb != 1 || o->which != 0) return 0;
  ob *B = &c->b[0];
  int nd = o->nd, nt = nd - 1;
  int cand[3], nc = 0;
  for (int e = 0; e < nd; e++) {
    int a = o->ax[e];
    int n = B->n[a];
    double h = B->df[a][1], dev = 0;
    for (int q = 0; q <= n + 1; q++) { dev = fmax(dev, fabs(B->df[a][q] - h)); dev = fmax(dev, fabs(B->dc[a][q] - h)); }
    char bc2[2] = {B->cbcpre[a][0], B->cbcpre[a][1]};
    int kf, kb; double nrm[2];
    if (dev <= 1e-12 * h && oc_find_fft(bc2, 1, &kf, &kb, nrm) == 0) cand[nc++] = e;
  }
  if (nc < nt) return 0;
  /* the line axis: a non-candidate if any, else the last candidate; it must
   * not be periodic                                                         */
  int ls = -1;
  for (int e = nd - 1; e >= 0; e--) {
    int isc = 0; for (int q = 0; q < nc; q++) if (cand[q] == e) isc = 1;
    if (!isc) { ls = e; break; }
  }
  if (ls < 0) {
    for (int q = nc - 1; q >= 0; q--) if (c->periods[o->ax[cand[q]]] == 0) { ls = cand[q]; break; }
  }
  if (ls < 0 || c->periods[o->ax[ls]] != 0) return 0;
  int k = 0;
  for (int e = 0; e < nd; e++) if (e != ls) o->tt[k++] = e;
  for (int t = 0; t < nt; t++) {
    int e = o->tt[t]; int isc = 0;
    for (int q = 0; q < nc; q++) if (cand[q] == e) isc = 1;
    if (!isc) return 0;
  }
  o->nt = nt; o->ts = ls;
  for (int t = 0; t < nt; t++) {
    int a = o->ax[o->tt[t]], n = B->n[a];
    char bc2[2] = {B->cbcpre[a][0], B->cbcpre[a][1]};
    double nrm[2];
    oc_find_fft(bc2, 1, &o->t_kind_fwd[t], &o->t_kind_bwd[t], nrm);
    o->t_norm[t] = 1.0 / (nrm[0] * (n + nrm[1]));
    o->t_lambda[t] = xcalloc(n, sizeof(double));
    oc_eigenvalues(n, bc2, 1, o->t_lambda[t]);
    o->th[t] = B->df[a][1];
    for (int q = 0; q < n; q++) o->t_lambda[t][q] /= o->th[t] * o->th[t];
  }
  o->tensor_ok = 1;
  return 1;
}

static int solve_tensor(const oc_case *c, oop *o, int mode, const double *b, double *x) {
  const ob *B = &c->b[0];
  int nd = o->nd;
  int nn[3] = {1, 1, 1};
  for (int e = 0; e < nd; e++) nn[e] = B->n[o->ax[e]];
  long st[3] = {1, nn[0], (long)nn[0] * nn[1]};
  long n = o->nunk;
  double *y = xcalloc(n, sizeof(double));
  memcpy(y, b, sizeof(double) * n);
  int maxn = 0; for (int e = 0; e < 3; e++) if (nn[e] > maxn) maxn = nn[e];
  double *lin = xcalloc(maxn, sizeof(double)), *lout = xcalloc(maxn, sizeof(double));
  /* forward transforms along the uniform axes */
  for (int t = 0; t < o->nt; t++) {
    int e = o->tt[t], o1 = -1, o2 = -1;
    for (int q = 0; q < 3; q++) if (q != e) { if (o1 < 0) o1 = q; else o2 = q; }
    for (int j2 = 0; j2 < nn[o2]; j2++)
      for (int j1 = 0; j1 < nn[o1]; j1++) {
        long base = j1 * st[o1] + j2 * st[o2];
        for (int q = 0; q < nn[e]; q++) lin[q] = y[base + q * st[e]];
        oc_r2r(o->t_kind_fwd[t], nn[e], lin, lout);
        for (int q = 0; q < nn[e]; q++) y[base + q * st[e]] = lout[q];
      }
  }
  /* tridiagonal solves along the line axis; the line-axis part of the
   * operator is rebuilt from the metrics exactly as the assembly does, the
   * transformed axes enter through their eigenvalues times the cell weight */
  int e = o->ts, a = o->ax[e], m = nn[e], o1 = -1, o2 = -1;
  for (int q = 0; q < 3; q++) if (q != e) { if (o1 < 0) o1 = q; else o2 = q; }
  double sg0 = B->cbcpre[a][0] == 'N' ? 1.0 : (B->cbcpre[a][0] == 'D' ? -1.0 : 0.0);
  double sg1 = B->cbcpre[a][1] == 'N' ? 1.0 : (B->cbcpre[a][1] == 'D' ? -1.0 : 0.0);
  double *cp = xcalloc(m, sizeof(double)), *dp = xcalloc(m, sizeof(double));
  for (int j2 = 0; j2 < nn[o2]; j2++)
    for (int j1 = 0; j1 < nn[o1]; j1++) {
      long base = j1 * st[o1] + j2 * st[o2];
      int jj[3] = {0, 0, 0}; jj[o1] = j1; jj[o2] = j2;
      double lsum = op_lam(o, mode);
      for (int t = 0; t < o->nt; t++) lsum += o->t_lambda[t][jj[o->tt[t]]];
      int sing = !o->has_dirichlet && lsum == 0.0;
      int mm = sing ? m - 1 : m; /* singular line: pin the last unknown */
      for (int q = 0; q < mm; q++) {
        long cell = base + q * st[e];
        int idx[3]; for (int oo = 0; oo < 3; oo++) idx[oo] = (oo == e) ? q : jj[oo];
        double other = 1.0;
        for (int oo = 0; oo < nd; oo++) if (oo != e) other *= B->df[o->ax[oo]][idx[oo] + 1];
        double rm = 1.0 / B->dc[a][q] * other, rp = 1.0 / B->dc[a][q + 1] * other;
        double lo = rm, up = rp, dg = -(rm + rp);
        if (c->periods[a] == 0) {
          if (B->is_bound[0][a] && q == 0) { dg += sg0 * rm; lo = 0.0; }
          if (B->is_bound[1][a] && q == m - 1) { dg += sg1 * rp; up = 0.0; }
        }
        dg += lsum * o->wgt[cell];
        double rhs = y[cell];
        if (q == 0) { cp[0] = up / dg; dp[0] = rhs / dg; }
        else { double den = dg - lo * cp[q - 1]; cp[q] = up / den; dp[q] = (rhs - lo * dp[q - 1]) / den; }
      }
      if (mm < m) y[base + (m - 1) * st[e]] = 0.0;
      if (mm == 0) continue;
      double xv = dp[mm - 1];
      y[base + (mm - 1) * st[e]] = xv;
      for (int q = mm - 2; q >= 0; q--) { xv = dp[q] - cp[q] * xv; y[base + q * st[e]] = xv; }
    }
  free(cp); free(dp);
  /* inverse transforms */
  for (int t = o->nt - 1; t >= 0; t--) {
    int et = o->tt[t], p1 = -1, p2 = -1;
    for (int q = 0; q < 3; q++) if (q != et) { if (p1 < 0) p1 = q; else p2 = q; }
    for (int j2 = 0; j2 < nn[p2]; j2++)
      for (int j1 = 0; j1 < nn[p1]; j1++) {
        long base = j1 * st[p1] + j2 * st[p2];
        for (int q = 0; q < nn[et]; q++) lin[q] = y[base + q * st[et]];
        oc_r2r(o->t_kind_bwd[t], nn[et], lin, lout);
        for (int q = 0; q < nn[et]; q++) y[base + q * st[et]] = lout[q] * o->t_norm[t];
      }
  }
  memcpy(x, y, sizeof(double) * n);
  free(y); free(lin); free(lout);
  return 0;
}

int oc_set_solver(oc_case *c, int method, double tol, int maxiter) {
  c->method = method;
  if (tol > 0) c->tol = tol;
  if (maxiter > 0) c->maxiter = maxiter;
  return 0;
}

static int pick_method(oc_case *c, oop *o) {
  int m = c->method;
  if (m == OC_SOLVE_AUTO) {
    if (tensor_setup(c, o)) return OC_SOLVE_TENSOR;
    if (band_setup(c, o)) return OC_SOLVE_BAND;
    return OC_SOLVE_CG;
  }
  if (m == OC_SOLVE_TENSOR && !tensor_setup(c, o)) return -1;
  if (m == OC_SOLVE_BAND && !band_setup(c, o)) return -1;
  return m;
}

static int solve_op(oc_case *c, oop *o, int mode, const double *b, double *x) {
  if (mode < 0 || mode >= o->nmodes) return -1;
  int m = pick_method(c, o);
  switch (m) {
  case OC_SOLVE_TENSOR: return solve_tensor(c, o, mode, b, x);
  case OC_SOLVE_BAND: return solve_band(o, mode, b, x);
  case OC_SOLVE_CG: return solve_cg(c, o, mode, b, x) >= c->maxiter ? -2 : 0;
  default: return -1;
  }
}

int oc_solve_operator(oc_case *c, int mode, const double *b, double *x) { return solve_op(c, &c->P, mode, b, x); }

/* gather/scatter the unknowns of one mode plane (2D) or of the whole
 * interior (3D) of a ghosted field                                          */
static void gather(const oc_case *c, const oop *o, const double *const *F, int mode, double *vec) {
  int d = c->fft_axis - 1;
  for (int ib = 0; ib < c->nb; ib++) {
    const ob *B = &c->b[ib];
    const int *e = o->ext[ib];
    long q = o->off[ib];
    if (o->nd == 3) {
      for (int k = 1; k <= e[2]; k++)
        for (int j = 1; j <= e[1]; j++)
          for (int i = 1; i <= e[0]; i++) vec[q++] = F[ib][IDX(B, i, j, k)];
    } else {
      int a0 = o->ax[0], a1 = o->ax[1];
      for (int i2 = 1; i2 <= e[a1]; i2++)
        for (int i1 = 1; i1 <= e[a0]; i1++) {
          int ijk[3]; ijk[d] = mode + 1; ijk[a0] = i1; ijk[a1] = i2;
          vec[q++] = F[ib][IDX(B, ijk[0], ijk[1], ijk[2])];
        }
    }
  }
}
static void scatter(const oc_case *c, const oop *o, double *const *F, int mode, const double *vec) {
  int d = c->fft_axis - 1;
  for (int ib = 0; ib < c->nb; ib++) {
    const ob *B = &c->b[ib];
    const int *e = o->ext[ib];
    long q = o->off[ib];
    if (o->nd == 3) {
      for (int k = 1; k <= e[2]; k++)
        for (int j = 1; j <= e[1]; j++)
          for (int i = 1; i <= e[0]; i++) F[ib][IDX(B, i, j, k)] = vec[q++];
    } else {
      int a0 = o->ax[0], a1 = o->ax[1];
      for (int i2 = 1; i2 <= e[a1]; i2++)
        for (int i1 = 1; i1 <= e[a0]; i1++) {
          int ijk[3]; ijk[d] = mode + 1; ijk[a0] = i1; ijk[a1] = i2;
          F[ib][IDX(B, ijk[0], ijk[1], ijk[2])] = vec[q++];
        }
    }
  }
}

/* in-place r2r of length n along the FFT axis over the interior lines of a
 * field, every block; the result scaled                                    */
static void fft_axis_transform(oc_case *c, double *const *F, int kind, int n, double scale) {
  int d = c->fft_axis - 1;
  for (int ib = 0; ib < c->nb; ib++) {
    ob *B = &c->b[ib];
    int a1, a2; tang(d, &a1, &a2);
    double *G = F[ib];
#pragma omp parallel
    {
      double *lin = xcalloc(n, sizeof(double)), *lout = xcalloc(n, sizeof(double));
#pragma omp for collapse(2) schedule(static)
      for (int t2 = 1; t2 <= B->n[a2]; t2++)
        for (int t1 = 1; t1 <= B->n[a1]; t1++) {
          for (int q = 0; q < n; q++) lin[q] = G[plane_idx(B, d, q + 1, t1, t2)];
          oc_r2r(kind, n, lin, lout);
          for (int q = 0; q < n; q++) G[plane_idx(B, d, q + 1, t1, t2)] = lout[q] * scale;
        }
      free(lin); free(lout);
    }
  }
}

/* transform, per-mode solve, inverse transform (main.f90:1004-1022,
 * solve_helmholtz.f90:83-99); sol (may be NULL) is the warm start and
 * receives the solution (po / phio, solver.f90:909-920)                     */
static int solve_field(oc_case *c, oop *o, double *const *F, double *const *sol) {
  long n = o->nunk;
  double *b = xcalloc(n, sizeof(double)), *x = xcalloc(n, sizeof(double));
  int rc = 0;
  int nf = o->nd == 2 ? c->b[0].n[c->fft_axis - 1] - o->qfft : 0;
  if (o->nd == 2) fft_axis_transform(c, F, o->kind_fwd, nf, 1.0);
  for (int m = 0; m < o->nmodes && rc == 0; m++) {
    gather(c, o, (const double *const *)F, m, b);
    if (sol) gather(c, o, (const double *const *)sol, m, x);
    else memset(x, 0, sizeof(double) * n);
    rc = solve_op(c, o, m, b, x);
    scatter(c, o, F, m, x);
    if (sol) scatter(c, o, sol, m, x);
  }
  if (o->nd == 2) {
    fft_axis_transform(c, F, o->kind_bwd, nf, 1.0);
    /* the normalisation covers every interior point (main.f90:1020) */
    for (int ib = 0; ib < c->nb; ib++) {
      ob *B = &c->b[ib];
      for (int k = 1; k <= B->n[2]; k++)
        for (int j = 1; j <= B->n[1]; j++)
          for (int i = 1; i <= B->n[0]; i++) F[ib][IDX(B, i, j, k)] *= o->normfft;
    }
  }
  free(b); free(x);
  return rc;
}

int oc_solve_pressure(oc_case *c) {
  double *F[MAXB], *S[MAXB];
  for (int ib = 0; ib < c->nb; ib++) { F[ib] = c->b[ib].f[4]; S[ib] = c->b[ib].f[5]; }
  return solve_field(c, &c->P, F, S);
}

/* ------------------------------------------------------------------------ */
/* implicit diffusion (_IMPDIFF)                                             */
/* ------------------------------------------------------------------------ */
int oc_set_impdiff(oc_case *c, int on, char *err, int errlen) {
  for (int q = 0; q < 3 + MAXS; q++) free_op(&c->H[q]);
  c->impdiff = 0;
  if (!on) return 0;
  for (int q = 0; q < 3 + c->nscal; q++) {
    oop *o = &c->H[q];
    o->which = q + 1;
    o->stag = q < 3 ? q : -1;
    if (assemble_op(c, o, err, errlen)) {
      for (int r = 0; r <= q; r++) free_op(&c->H[r]);
      return -1;
    }
  }
  c->impdiff = 1;
  return 0;
}

/* solve_impdiff_field solve_helmholtz.f90:30-100 (_fft) / 101-136 (_3d):
 * phi = phi * alphai * dx2 (non-FFT axes), updt_rhs with the BC planes over
 * the unknowns, then the Helmholtz solve with alphai on the weighted
 * diagonal (add_weighted_constant_to_(n_)diagonal(s)).  which: 1..3 = u, v,
 * w, 4+i = scalar i; alpha = -visc*dtrk/2 (-alpha_s*dtrk/2 for a scalar).   */
int oc_solve_impdiff(oc_case *c, int which, double alpha) {
  if (!c->impdiff || which < 1 || which > 3 + c->nscal) return -1;
  oop *o = &c->H[which - 1];
  const double alphai = 1.0 / alpha;
  double *F[MAXB];
  for (int ib = 0; ib < c->nb; ib++) {
    ob *B = &c->b[ib];
    F[ib] = which <= 3 ? B->f[which - 1] : B->sv[which - 4];
    double *G = F[ib];
    for (int k = 1; k <= B->n[2]; k++)
      for (int j = 1; j <= B->n[1]; j++)
        for (int i = 1; i <= B->n[0]; i++) {
          int ijk[3] = {i, j, k};
          double v = G[IDX(B, i, j, k)] * alphai;
          for (int a = 0; a < 3; a++)
            if (a != c->fft_axis - 1) v = v * wdx2(o, B, a, ijk[a]);
          G[IDX(B, i, j, k)] = v;
        }
    /* updt_rhs bound.f90:543-585 with hi = the unknowns' upper bound */
    for (int d = 0; d < 3; d++)
      for (int s = 0; s < 2; s++) {
        if (!B->is_bound[s][d] || !o->rhs[ib][d][s]) continue;
        int a1, a2; tang(d, &a1, &a2);
        int q = s == 0 ? 1 : o->ext[ib][d];
        long m1 = o->ext[ib][a1], m2 = o->ext[ib][a2];
        for (int t2 = 1; t2 <= m2; t2++)
          for (int t1 = 1; t1 <= m1; t1++)
            G[plane_idx(B, d, q, (int)t1, (int)t2)] += o->rhs[ib][d][s][(t1 - 1) + m1 * (t2 - 1)];
      }
  }
  o->shift = alphai;
  int rc = solve_field(c, o, F, NULL);
  o->shift = 0.0;
  return rc;
}

void oc_init_bounds(oc_case *c) {
  oc_bounduvw(c, 0);
  oc_boundp(c, 3);
  for (int is = 0; is < c->nscal; is++) oc_boundp(c, 9 + 2 * is);
  oc_inflow(c, 0);
}

/* main.f90:893-1033, one time step                                          */
int oc_step(oc_case *c, double dt) {
  for (int irk = 1; irk <= 3; irk++) {
    double dtrk = (RKC[irk - 1][0] + RKC[irk - 1][1]) * dt;
    if (c->nscal > 0) {
      oc_rk_scal(c, irk, dt, NULL);
      for (int is = 0; is < c->nscal; is++) {
        if (c->impdiff) { /* main.f90:914-936 */
          int rc = oc_solve_impdiff(c, 4 + is, -c->salpha[is] * dtrk / 2.0);
          if (rc) return rc;
        }
        oc_boundp(c, 9 + 2 * is);
      }
    }
    const double alpha = -c->visc * dtrk / 2.0;
    oc_rk_mom(c, irk, dt);
    if (c->impdiff) /* main.f90:947-997 */
      for (int q = 1; q <= 3; q++) {
        int rc = oc_solve_impdiff(c, q, alpha);
        if (rc) return rc;
      }
    oc_bounduvw(c, 0);
    oc_inflow(c, 0);
    oc_fillps(c, dtrk);
    int rc = oc_solve_pressure(c);
    if (rc) return rc;
    oc_boundp(c, 4);
    oc_correc(c, dtrk);
    oc_bounduvw(c, 1);
    oc_inflow(c, 1);
    if (c->impdiff) oc_updt_pressure_impdiff(c, alpha);
    else oc_updt_pressure(c);
    oc_boundp(c, 3);
  }
  return 0;
}
