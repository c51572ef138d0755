// Shared declarations of the gfx950 kernels and their host launchers.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdexcept>
#include <string>

namespace snac {

#define SNAC_HIP(x)                                                                              \
  do {                                                                                           \
    hipError_t e_ = (x);                                                                         \
    if (e_ != hipSuccess)                                                                        \
      throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e_) + " (" + #x + \
                               ") at " + __FILE__ + ":" + std::to_string(__LINE__));            \
  } while (0)

// Extents of one block: interior sizes and strides of the ghosted arrays
// (column-major, i fastest, index 0 = the lo-1 ghost).
struct Geo {
  int nx, ny, nz;
  long sx, sy;  // sx = nx+2, sy = ny+2
};

// 1D grid metrics of one block on the device (n+2 entries each), and their
// reciprocals: the kernels multiply by 1/d instead of dividing by d, which
// removes the FP64 division sequences from the stencils (a rounding-level
// difference from the reference's divisions).
struct Metrics {
  const double *dxc, *dxf, *dyc, *dyf, *dzc, *dzf;
  const double *idxc, *idxf, *idyc, *idyf, *idzc, *idzf;
};

// One ghost-plane operation: dst plane (normal d, index qdst) =
//   c0 + c1 * src plane (index qsrc, same tangential element) + c2 * arr[t]
// over the full ghosted tangential extent m1 x m2.  Covers halo copies
// (updthalo), every set_bc form and the inflow planes.
struct FaceOp {
  double *dst;
  const double *src;
  const double *arr;
  long dsx, dsy, ssx, ssy;  // strides of dst / src arrays
  int d, qdst, qsrc, m1, m2;
  double c0, c1, c2;
};

// A ghost-fill program composed down to its final values: each ghost cell
// the program writes is one chain  v = base cell (or 0); for each level
// v = c0 + c1 * v + c2 * arr[a0 + i]  -- the ops that produced the cell in
// program order, evaluated in that order (so the values are the phased
// program's bit for bit).  A segment is a run of such cells along one row of
// one op's plane with the same chain shape and constant index steps.
constexpr int SEG_LEV = 4;
struct CellLev {
  double c0, c1, c2;
  const double *arr;  // null: no array term
  long a0;            // arr index of the segment's first cell (step 1)
  int src;            // 0: the level does not read the previous value
};
struct CellSeg {
  double *dst;
  long d0, dstr;
  const double *base;  // null: the chain starts from a level with src == 0
  long b0, bstr;
  int len, nlev;
  CellLev lev[SEG_LEV];
};
void launch_cell_segs(hipStream_t s, const CellSeg *dev_segs, int nseg);

// ---- stencil kernels (kernels_stencil.hip) ----
// rk_mom: reads u,v,w,p (ghosted) and dudtrko; writes the updated interior
// velocities to un,vn,wn and dudtrk to dto.  The three tendencies share one
// array of 3 nx ny nz doubles laid out [k][j][component][i]: dtv = dtu + nx,
// dtw = dtu + 2 nx, rows 3 nx apart.
// s1 (scalar 1, ghosted) non-null adds the Boussinesq term with gb = gacc*beta;
// variant 0: one cell per thread; 1..6: k-march over chunks of 8/16/32
// planes with 64x4 or 32x8 tiles; 7..10: k-march with LDS-staged planes,
// 64x4 / 64x8 tiles over 16 / 32 planes; 11..14: 7..10 with non-temporal
// hints on the read-once / write-once streams; 15, 16: 64x8 tiles over 64
// planes without / with the hints (autotuned).  impdiff: the diffusion
// is kept out of dudtrk and half of it stays in the update (the reference's
// _IMPDIFF build, rk.f90:29-127); the Helmholtz solves add the other half.
void launch_rk_mom(hipStream_t s, const Geo &g, const Metrics &m, const double *u, const double *v,
                   const double *w, const double *p, double *un, double *vn, double *wn, double *dtu,
                   double *dtv, double *dtw, double f1, double f2, double f12, double visc, double bfx,
                   double bfy, double bfz, const double *s1, const double gb[3], int variant, bool impdiff);
// Mixed precision: the same substage with the right-hand side evaluated in
// FP32 from FP32 copies of u, v, w staged in LDS and the old tendencies
// stored in FP32 (same [k][j][component][i] layout, in floats); the RK3
// accumulation, the pressure gradient and the body forces stay FP64.  LDS
// variants 7..16 only.
void launch_rk_mom_f32(hipStream_t s, const Geo &g, const Metrics &m, const double *u, const double *v,
                       const double *w, const double *p, double *un, double *vn, double *wn, float *dtu, float *dtv,
                       float *dtw, double f1, double f2, double f12, double visc, double bfx, double bfy, double bfz,
                       const double *s1, const double gb[3], int variant, bool impdiff);
// element-wise conversions between FP64 and FP32 arrays of n values
void launch_cvt_d2f(hipStream_t s, const double *a, float *b, long n);
void launch_cvt_f2d(hipStream_t s, const float *a, double *b, long n);
// rk_scal of one scalar: reads sc (ghosted), writes the interior of sn and dso.
// imp (implicit diffusion, rk.f90:151-202): 0 off; 1 the diffusion out of
// dsdtrk, half of it left in sn; 2 likewise but the half is taken off by
// launch_bulk_force after the bulk mean, the diffusion written to ddo
// (interior layout)
void launch_rk_scal(hipStream_t s, const Geo &g, const Metrics &m, const double *u, const double *v, const double *w,
                    const double *sc, double *sn, double *dso, double alpha, double f1, double f2, double f12,
                    double src, int imp = 0, double *ddo = nullptr);
// partial volume integrals of f (reduce_nparts(g) entries), and the bulk
// forcing f += scalf - red[0]*ivol (fout, if given, receives that value);
// dd non-null first subtracts h*dd (interior layout)
void launch_volsum_part(hipStream_t s, const Geo &g, const Metrics &m, const double *f, double *psum);
void launch_bulk_force(hipStream_t s, const Geo &g, double *f, const double *red, double scalf, double ivol,
                       double *fout, const double *dd = nullptr, double h = 0.0);
// fillps with updt_rhs fused; rhs planes may be null (ordered x0,x1,y0,y1,z0,z1)
void launch_fillps(hipStream_t s, const Geo &g, const Metrics &m, int fft_axis, const double *u,
                   const double *v, const double *w, double *pp, double dtrk, const double *const rhsp[6]);
// correc out-of-place over the whole ghosted array (copy outside the
// reference's loop ranges); optionally fuses updt_pressure (p += pp).
void launch_correc(hipStream_t s, const Geo &g, const Metrics &m, const double *pp, const double *u,
                   const double *v, const double *w, double *un, double *vn, double *wn, double dtrk,
                   double *p_update);
void launch_updt_pressure(hipStream_t s, const Geo &g, const double *pp, double *p);
// reductions: partial results into part (size >= nparts(g)); final on device
int reduce_nparts(const Geo &g);
void launch_chkdiv_part(hipStream_t s, const Geo &g, const Metrics &m, const double *u, const double *v,
                        const double *w, double vol, double *part_sum, double *part_max);
void launch_chkdt_part(hipStream_t s, const Geo &g, const Metrics &m, const double *u, const double *v,
                       const double *w, double *part_max);
// reduces n partial sums and maxima into out[0] (sum) and out[1] (max)
void launch_reduce_final(hipStream_t s, const double *part_sum, const double *part_max, int n, double *out);

// ---- implicit diffusion (kernels_impdiff.hip) ----
// Right-hand side of one field's Helmholtz problem over a block: the
// unknowns' extents (the block's, less the bounded upper face along a
// velocity component's own axis), per-axis weight factors (dx2 of the
// non-FFT axes, 1 along the FFT axis; n entries each, device) and the BC
// planes of the bounded faces over the unknowns (first tangential axis
// fastest; null where absent).
struct ImpRhs {
  int ext[3];
  const double *w[3];
  const double *plane[3][2];
};
// out (ghosted, same layout as phi) = phi*alphai*w0*w1*w2 + BC planes over
// the unknowns, 0 past them
void launch_imp_rhs(hipStream_t s, const Geo &g, const ImpRhs &r, const double *phi, double *out, double alphai);
// dst = src over the unknowns (indices 1..ext of each axis) of a block
void launch_copy_unknowns(hipStream_t s, const Geo &g, const int ext[3], const double *src, double *dst);
// updt_pressure of the _IMPDIFF build: p += pp + alpha * Laplacian(pp)
void launch_updt_pressure_imp(hipStream_t s, const Geo &g, const Metrics &m, const double *pp, double *p,
                              double alpha);

// plain device copy of n doubles (n even), the HBM reference kernel
void launch_stream_copy(hipStream_t s, const double *a, double *b, long n);

// ---- ghost planes (kernels_bound.hip) ----
void launch_face_ops(hipStream_t s, const FaceOp *dev_ops, int nops, int max_plane);

// ---- transforms (kernels_r2r.hip) ----
// Batched r2r along lines: line (b1,b2) element q at in[q*ies + b1*is1 + b2*is2],
// written to out[q*oes + b1*os1 + b2*os2] times scale.  tw: per-size table
// from r2r_table(); mat: dense n x n matrix (row = output) or null.
struct R2RPlan {
  int n = 0, kind = 0;
  bool fft = false;        // LDS FFT path (n a power of two >= 4, supported kind)
  int lshift = 0;          // FFT tile (autotuned): 0..2 lines per workgroup = max >> lshift;
                           // 3, 4 (R2HC/HC2R): full tiles, 2 / 4 per workgroup, pipelined
                           // 5 (R2HC/HC2R): full tiles, wavefront-local passes;
                           // 6 (R2HC/HC2R): full tiles, the FFT's factors swapped;
                           // 7 (R2HC/HC2R): wavefront-local passes, factors swapped
  const double *tab = nullptr;  // device twiddles (fft) or dense matrix
};
R2RPlan r2r_plan(int kind, int n);  // cached per (kind, n); throws if unsupported
int r2r_tile_choices(const R2RPlan &pl);  // number of tile choices of the plan's kernel
void launch_r2r(hipStream_t s, const R2RPlan &pl, const double *in, long ies, long is1, long is2,
                double *out, long oes, long os1, long os2, int nb1, int nb2, double scale);
// host reference of FFTW's r2r definitions, used to build dense matrices
double r2r_coef(int kind, int n, int k, int j);

// ---- tridiagonal line solves (kernels_tridiag.hip) ----
// Lines along axis s of a (G0,G1,G2) array (i fastest).  For the line at
// transformed indices (t0,t1) of the other two axes:
//   lam = lamA[idxA] + lamB[idxB] + shift ; diag[g] = d0[g] + lam * wt[g]
// (shift: 1/alpha of an implicit-diffusion Helmholtz solve, 0 for the pressure)
//   lower[g] = am[g], upper[g] = ap[g]
// idxA/idxB are the indices along the two non-line axes (lamA/lamB sized to
// them).  Lines with lam == 0 and `pin_when_zero` get their last unknown
// fixed to zero (singular Neumann-Neumann line).
struct TridiagLine {
  int G[3];
  int s;        // line axis
  int variant;  // 0: c' in LDS (lines up to 128), 1: c' in cscratch, 2: chunked, no scratch,
                // 3: chunked with twice the chunks (autotuned)
  const double *am, *ap, *d0, *wt;  // size G[s]
  const double *lamA, *lamB;        // along the other axes (ascending order)
  int pin_when_zero;
  double shift = 0.0;
};
// dscratch (d') and cscratch (c', used only for lines longer than the LDS
// budget) have the size and layout of x
void launch_tridiag(hipStream_t st, const TridiagLine &t, double *x, double *dscratch, double *cscratch);
// chunks per line of variant 2 (fine: variant 3) for lines of length m
// (0: not available)
int tridiag_chunks(int m, bool fine = false);

}  // namespace snac
