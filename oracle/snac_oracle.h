/*
 * snac_oracle.h -- CPU FP64 restatement of SNaC's time-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path in
 * snac_amd/: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it.  The product library never links or calls it.
 *
 * Every routine restates one reference routine; the citation is given next to
 * each declaration (paths relative to the SNaC source tree).
 *
 * Conventions (same as the reference): column-major, i fastest; every 3D field
 * of a block carries one ghost layer, so local index i in [0, n+1] maps to the
 * Fortran index lo-1+i.  Metric arrays (dxc,dxf,...) have n+2 entries with the
 * same offset.
 *
 * Parity status per function is summarised in the header of snac_oracle.c.
 */
#ifndef SNAC_ORACLE_H
#define SNAC_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* r2r transform kinds, numbered as FFTW's (fftw.f90:31-54) */
enum { OC_R2HC = 0, OC_HC2R = 1, OC_REDFT00 = 3, OC_REDFT01 = 4, OC_REDFT10 = 5,
       OC_REDFT11 = 6, OC_RODFT00 = 7, OC_RODFT01 = 8, OC_RODFT10 = 9, OC_RODFT11 = 10 };

/* Poisson solver used by the oracle for the pressure correction */
enum { OC_SOLVE_BAND = 0, OC_SOLVE_CG = 1, OC_SOLVE_TENSOR = 2, OC_SOLVE_AUTO = 3 };

typedef struct oc_case oc_case;

/* ---- case set-up: param.f90:68-277 (inputs), initmpi.f90:9-254 (placement,
 *      neighbours, periods), initgrid.f90:9-127 (grid), main.f90:226-886 ----
 * Flat array layouts follow the Fortran namelist ordering:
 *   ng,lmin,lmax,gt,gr     : [b][dir]                     (3 per block)
 *   cbcvel,bcvel           : [b][comp][dir][side]         (18 per block)
 *   cbcpre,bcpre,inflow    : [b][dir][side]               (6 per block)
 * fft_axis: 0 = full 3D Poisson, 1..3 = FFT along x/y/z.                    */
oc_case *oc_create(int nblocks, const int *ng, const double *lmin, const double *lmax,
                   const int *gt, const double *gr,
                   const char *cbcvel, const double *bcvel,
                   const char *cbcpre, const double *bcpre,
                   const int *inflow_type,
                   int fft_axis, double visci, const double *bforce,
                   char *errbuf, int errlen);
void oc_destroy(oc_case *c);
int  oc_set_solver(oc_case *c, int method, double tol, int maxiter);
int  oc_set_threads(int nthreads);

/* geometry queries (for tests) */
int  oc_nblocks(const oc_case *c);
void oc_block_info(const oc_case *c, int b, int *n3, int *lo_g3, int *nb6, int *periods3);
double oc_vol_all(const oc_case *c);
/* which: 0..5 = dxc,dxf,dyc,dyf,dzc,dzf ; 6..11 = xc,xf,yc,yf,zc,zf (n+2 entries) */
int  oc_get_metric(const oc_case *c, int b, int which, double *out);

/* field access; fid: 0 u,1 v,2 w,3 p,4 pp,5 po (ghosted (n+2)^3),
 *                    6 dudtrko,7 dvdtrko,8 dwdtrko (interior n^3),
 *                    9+2i scalar i (ghosted), 10+2i its dsdtrko (interior) */
int  oc_field_size(const oc_case *c, int b, int fid);
int  oc_get_field(const oc_case *c, int b, int fid, double *out);
int  oc_set_field(oc_case *c, int b, int fid, const double *in);

/* scalar transport inputs (param.f90:96-100,148-163): per-scalar alphai,
 * ssource, is_sforced, scalf; beta and gacc(3) of the Boussinesq term;
 * cbcscal/bcscal flat [b][iscal][dir][side].                                */
int oc_set_scalars(oc_case *c, int nscal, const double *alphai, double beta, const double *gacc,
                   const double *ssource, const int *is_sforced, const double *scalf,
                   const char *cbcscal, const double *bcscal);
/* the _BOUSSINESQ_BUOYANCY build (rk.f90:104-108) as a switch, off by
 * default; on needs nscal > 0 (param.f90:166-170): returns -1 otherwise   */
int oc_set_buoyancy(oc_case *c, int enable);

/* ---- hot-path routines (all blocks at once) ---- */
/* rk_scal rk.f90:130-204 + bulk_forcing_s scal.f90:182-199 for every scalar;
 * f (nscal entries, may be NULL) receives the bulk forcing of this substage */
void oc_rk_scal(oc_case *c, int irk, double dt, double *f);
/* rk.f90:7-128 + mom.f90:112-386, rkcoeff param.f90:11-13; irk in 1..3      */
void oc_rk_mom(oc_case *c, int irk, double dt);
/* bound.f90:8-66 (+ updthalo 587-650 over the block neighbour table)        */
void oc_bounduvw(oc_case *c, int is_correc);
/* bound.f90:68-106 ; which: 3 = p, 4 = pp, 9+2i = scalar i (its own BCs)    */
void oc_boundp(oc_case *c, int which);
/* bound.f90:438-477                                                          */
void oc_inflow(oc_case *c, int is_correc);
/* fillps.f90:7-54 followed by updt_rhs bound.f90:543-585                     */
void oc_fillps(oc_case *c, double dtrk);
/* fft.f90:85-94 fwd, solver.f90:843-922 / 524-573, fft bwd, main.f90:1017-1020 */
int  oc_solve_pressure(oc_case *c);
/* correc.f90:7-48                                                            */
void oc_correc(oc_case *c, double dtrk);
/* updt_pressure.f90:23-29                                                    */
void oc_updt_pressure(oc_case *c);
/* updt_pressure.f90:23-45 of the _IMPDIFF build (alpha = -visc*dtrk/2)       */
void oc_updt_pressure_impdiff(oc_case *c, double alpha);
/* chkdiv.f90:8-44                                                            */
void oc_chkdiv(const oc_case *c, double *divtot, double *divmax);
/* chkdt.f90:8-56                                                             */
double oc_chkdt(const oc_case *c);
/* one full time step, main.f90:893-1033                                      */
int  oc_step(oc_case *c, double dt);
/* initial boundary fill as main.f90:562-639 (bounduvw, boundp, inflow)       */
void oc_init_bounds(oc_case *c);

/* ---- implicit diffusion (the reference's _IMPDIFF build) ----
 * oc_set_impdiff assembles the Helmholtz operators of u, v, w and of every
 * scalar (main.f90:769-886; call after oc_set_scalars) and switches rk_mom,
 * rk_scal, updt_pressure, chkdt and oc_step to their _IMPDIFF forms; an
 * unsupported case returns -1 with a message in err.  oc_solve_impdiff is
 * solve_impdiff_field (solve_helmholtz.f90:30-136) of one field: which 1..3 =
 * u, v, w, 4+i = scalar i; alpha = -visc*dtrk/2 (-alpha_s*dtrk/2).          */
int  oc_set_impdiff(oc_case *c, int on, char *err, int errlen);
int  oc_solve_impdiff(oc_case *c, int which, double alpha);

/* ---- transforms and eigenvalues (fft.f90:96-189) ---- */
/* direct O(n^2) definitions of FFTW's r2r kinds, on one contiguous line      */
int  oc_r2r_direct(int kind, int n, const double *in, double *out);
/* the same through the oracle's radix-2 path (n power of two), else direct   */
int  oc_r2r(int kind, int n, const double *in, double *out);
/* find_fft fft.f90:96-149: pressure (centered) kinds and normalisation       */
int  oc_find_fft(const char *bc2, int is_centered, int *kind_fwd, int *kind_bwd, double *norm2);
/* eigenvalues fft.f90:151-189 (DN added: same spectrum as ND)               */
int  oc_eigenvalues(int n, const char *bc2, int is_centered, double *lambda);
/* initgrid.f90:9-83 for one direction: outputs n+2 entries each              */
int  oc_initgrid(int n, int gt, double gr, double lmin, double lmax,
                 double *drc, double *drf, double *rc, double *rf);

/* ---- Poisson operator introspection (tests) ----
 * Applies the assembled operator of solver.f90 (2D per mode, or 3D) to a
 * vector laid out block after block (interior cells, i fastest); mode is the
 * 0-based index along the FFT axis (ignored for the 3D operator).           */
int  oc_num_unknowns(const oc_case *c);
int  oc_apply_operator(const oc_case *c, int mode, const double *x, double *y);
int  oc_solve_operator(oc_case *c, int mode, const double *b, double *x);
int  oc_get_lambda(const oc_case *c, double *lambda, double *normfft);

/* ---- the learned preconditioner's network (net_oracle.c; no reference
 *      counterpart): forward and loss gradient of the encoder-decoder of
 *      snac_amd/csrc/net.hpp.  x: B planes of 2 channels ([b][y][x][2]),
 *      t and y: B planes; H, W multiples of 8; parameters flattened per
 *      layer (weights [tap][cout][cin], then biases).  round = 1 rounds the
 *      convolution operands to BF16 as the MFMA kernels do. ---- */
long on_num_params(void);
int  on_forward(int B, int H, int W, const float *params, const float *x, float *y, int round);
int  on_loss_grad(int B, int H, int W, const float *params, const float *x, const float *t, float *loss,
                  float *grad, int round);

#ifdef __cplusplus
}
#endif
#endif
