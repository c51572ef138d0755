/*
 * snac_amd.h -- C-ABI of the MI355X (gfx950) time-step hot path of SNaC.
 *
 * A Fortran host binds these with iso_c_binding (the bind(C) interface module
 * is given in INTEGRATION.md); Python binds them with ctypes (snac_amd/).
 * Plain pointers and sizes only.  Every function returns 0 on success and a
 * negative code on failure; snac_last_error() then holds the message (the
 * Fortran shim turns a non-zero status into `error stop`, which is how the
 * reference reports failures).
 *
 * Threading: one host thread per context; calls on one context are not
 * re-entrant.  All fields live in device memory owned by the context; host
 * arrays are touched only by snac_upload_field / snac_download_field.
 *
 * Array layout at the boundary (identical to the reference's arrays,
 * main.f90:226-236): column-major, i fastest, one ghost layer per side, i.e.
 * a block with interior n = (nx,ny,nz) exchanges (nx+2)*(ny+2)*(nz+2) doubles
 * for u,v,w,p,pp,po and the scalars, nx*ny*nz doubles for dudtrko/dvdtrko/
 * dwdtrko and the scalars' dsdtrko.
 *
 * Reference interfaces replaced (paths relative to the SNaC source tree):
 * each declaration names the routine(s) it stands in for.
 */
#ifndef SNAC_AMD_H
#define SNAC_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct snac_ctx snac_ctx;

/* field ids for upload/download; scalar i (0-based, scalars(i+1)%val of
 * scal.f90:22-49) is SNAC_SCAL0 + 2*i, its dsdtrko SNAC_SCAL0 + 2*i + 1 */
enum {
  SNAC_U = 0, SNAC_V = 1, SNAC_W = 2, SNAC_P = 3, SNAC_PP = 4, SNAC_PO = 5,
  SNAC_DUDTRKO = 6, SNAC_DVDTRKO = 7, SNAC_DWDTRKO = 8, SNAC_SCAL0 = 16
};

/* Poisson solver choice for snac_ctx_set_solver: AUTO takes the direct
 * tensor-product solver when the union of blocks is a box with uniform
 * transform axes, else the batched MG-preconditioned CG (any union) */
enum { SNAC_SOLVER_AUTO = 0, SNAC_SOLVER_TENSOR = 1, SNAC_SOLVER_MG = 2 };

/* Message of the last failure on this host thread ("" if none). */
const char *snac_last_error(void);

/* Reads dns.nml (&dns, &hypre) and blocks.nml (&blocks) exactly as
 * read_input does (param.f90:68-277), places the blocks and finds their
 * neighbours (initmpi.f90:9-254, one piece per block), builds the grids
 * (initgrid.f90:9-127), the inflow planes (main.f90:600-636, init_inflow
 * initflow.f90:369-429), the pressure BC rhs planes (init_bc_rhs
 * solver.f90:44-157) and the Poisson set-up (init_fft_reduction
 * solver.f90:1124-1138, init_n_2d_matrices solver.f90:981-1020 or
 * init_matrix_3d solver.f90:158-372, create/setup_n_solvers 1058-1077).
 * fft_axis 0..3 stands in for the compile-time FFT_AXIS (flags-default.mk).
 * device: HIP device ordinal, or -1 for a host-only context (set-up and
 * queries only; every compute entry point then fails).  Fields start at 0. */
int snac_ctx_create_from_nml(snac_ctx **ctx, const char *dns_nml, const char *blocks_nml,
                             int fft_axis, int device);
/* finalize_n_matrices/fftend (main.f90:1144-1150) */
int snac_ctx_destroy(snac_ctx *ctx);

/* geometry and parameter queries */
int snac_ctx_num_blocks(const snac_ctx *ctx);
/* n3 = interior sizes, lo_g3 = global placement, nb6 = neighbour block per
 * (dir,side) as [x0,x1,y0,y1,z0,z1] (-1 = physical boundary), periods3     */
int snac_ctx_block_info(const snac_ctx *ctx, int b, int *n3, int *lo_g3, int *nb6, int *periods3);
/* which: 0..5 = dxc,dxf,dyc,dyf,dzc,dzf ; 6..11 = xc,xf,yc,yf,zc,zf (n+2) */
int snac_ctx_get_metric(const snac_ctx *ctx, int b, int which, double *out);
/* dparams: [visc, cfl, dtmax, dt_f, bforce(3), vol_all, time_max, tw_max,
 *           hypre_tol] (11); iparams: [nstep, icheck, nscal, hypre_maxiter,
 *           restart, stop_type(3)] (8)                                       */
int snac_ctx_get_params(const snac_ctx *ctx, double *dparams, int *iparams);
/* Poisson solver selection (the reference's &hypre choice has no effect on
 * the discrete solution at tight tolerance; the MG-PCG path stops each
 * mode at ||r|| <= hypre_tol ||b|| or after hypre_maxiter iterations, the
 * &hypre parameters of param.f90:101, solver.f90:389-451)                   */
int snac_ctx_set_solver(snac_ctx *ctx, int solver);
/* Arithmetic of the time step (no counterpart in the reference, whose
 * precision is the compile-time real(rp), types.f90:17-19):
 * SNAC_PREC_FP64 (the default) runs everything in FP64.  SNAC_PREC_MIXED
 * evaluates the momentum right-hand side of rk_mom (mom.f90:112-327) in
 * FP32 and keeps the old tendencies dudtrko/dvdtrko/dwdtrko in FP32, while
 * the RK3 accumulation, the pressure gradient and the body forces of the
 * update stay FP64 (the velocities are FP64 fields); on the MG-PCG path it
 * also runs the V-cycle (line smoothing, restriction, prolongation,
 * coarse solves) in FP32 under a flexible CG whose residuals, operator,
 * dot products and null-space removal stay FP64, so the solve still stops
 * at hypre_tol.  fillps, correc, updt_pressure, the ghost fills and the
 * direct tensor-product solver stay FP64 in every mode.
 * SNAC_PREC_LEARNED keeps the FP64 stencils and preconditions the same
 * flexible CG with the trained
 * network (snac_learned_train) on the FP32 residual of each mode, falling
 * back to the FP32 V-cycle for the rest of a solve when some mode's residual
 * reduction stalls; the FP64 residual still decides convergence (FFT path;
 * steps run eagerly).  Drops the captured step graph; the next graph step
 * re-measures the CG count. */
enum { SNAC_PREC_FP64 = 0, SNAC_PREC_MIXED = 1, SNAC_PREC_LEARNED = 2 };
int snac_set_precision(snac_ctx *ctx, int mode);
/* ---- the learned preconditioner in the solver (no counterpart in the
 *      reference).  collect: record the next solves' (right-hand side,
 *      solution) pairs of every mode, each scaled by the RMS of its
 *      right-hand side, up to npairs (MG-PCG path with an FFT axis; steps
 *      run eagerly while collecting).  train: Adam on the recorded pairs
 *      (the network is created on first use: 4 levels, 16-64 channels,
 *      BF16 MFMA, planes of the union's bounding box rounded up to 8);
 *      loss_per_epoch may be NULL.  options: the stall fallback (on by
 *      default: an iteration that leaves some mode's residual norm above
 *      0.5 of the previous one hands the rest of the solve to the V-cycle;
 *      off: the network preconditions every iteration) and the initial
 *      guess (off by default: on, each solve starts from the network's
 *      answer for its right-hand side).  stats:
 *      recorded pairs and solves that fell back to the V-cycle. ---- */
int snac_learned_collect(snac_ctx *ctx, int npairs);
int snac_learned_train(snac_ctx *ctx, int epochs, int batch, float lr, unsigned seed, float *loss_per_epoch);
int snac_learned_options(snac_ctx *ctx, int fallback, double stall_ratio, int initial_guess);
int snac_learned_stats(snac_ctx *ctx, int *pairs, int *fallbacks);
/* average time of one network forward over every mode (HIP events on the
 * context's stream, reps launches) and its convolution GFLOP             */
int snac_learned_time(snac_ctx *ctx, int reps, double *ms, double *gflop);
/* CG iterations of the last MG-PCG pressure solve and its largest relative
 * residual over the modes (HYPRE's StructPCGGetNumIterations /
 * GetFinalRelativeResidualNorm, solver.f90:553-569); synchronises.  Both
 * are 0 when the tensor-product solver ran.                                 */
int snac_solver_stats(snac_ctx *ctx, int *iters, double *max_relres);

/* Counters of MG-PCG pressure solves since the last call (then reset):
 * all solves; those whose captured iterations (graph steps) left some mode
 * above the tolerance, so that the host resumed them; those that stopped
 * at hypre_maxiter above the tolerance (HYPRE's StructPCGGetNumIterations
 * reaching maxiter, solver.f90:553-569).  No device work.                  */
int snac_solver_counters(snac_ctx *ctx, int *solves, int *continued, int *unconverged);
/* CG iterations each pressure solve runs inside a captured step graph on
 * the MG-PCG path before the host resumes it with convergence checks (the
 * solve still stops at hypre_tol either way).  0 (the default): one
 * host-checked step measures it (largest count + 1, raised when later
 * solves need more).  No counterpart in the reference.                      */
int snac_set_graph_cg_iters(snac_ctx *ctx, int n);

/* host <-> device copies of one field of one block (layout above)           */
int snac_upload_field(snac_ctx *ctx, int b, int fid, const double *host);
int snac_download_field(snac_ctx *ctx, int b, int fid, double *host);

/* ---- files in the reference's formats ----
 * load_one (load.f90:84-187): io 'w' writes block b's field fid as its
 * interior values (Fortran order, FP64) followed by [time, istep] (null
 * pointers write 0); io 'r' checks that the file holds exactly that many
 * doubles (the reference aborts on a mismatch; here the call fails with the
 * same message), sets the interior with zero ghosts and returns time and
 * nint(istep).                                                              */
int snac_load_one(snac_ctx *ctx, char io, const char *filename, int b, int fid, double *time, int *istep);
/* The checkpoint of main.f90:543-560 ('r') and 1090-1134 ('w', the
 * is_overwrite_save naming): <datadir>fld_b_NNN_{u,v,w,p}.bin and
 * fld_b_NNN_s_MMM.bin per scalar, NNN the 1-based block number.  After 'r',
 * po = p as main.f90:557; call snac_init_bounds next, as main.f90:562-639. */
int snac_restart_io(snac_ctx *ctx, char io, const char *datadir, double *time, int *istep);
/* write_visu_3d (output.f90:289-303): out3d writes the interior of the field
 * to <datadir><fname_bin> (the whole field: the reference ignores nskip,
 * output.f90:235-239) and write_log_output appends
 * '(A,A,A,9i5,E15.7,i10)' fname_bin varname nmin nmax nskip time istep to
 * <datadir><fname_log>, the input of the XDMF scripts.                     */
int snac_write_visu_3d(snac_ctx *ctx, const char *datadir, const char *fname_bin, const char *fname_log,
                       const char *varname, int b, int fid, const int *nmin, const int *nmax, const int *nskip,
                       double time, int istep);
/* save_grid (initgrid.f90:128-143) of every block and direction,
 * <datadir>grid_{x,y,z}_b_NNN.bin / .out, and geometry_b_NNN.out
 * (main.f90:429-437).  Works on host-only contexts.  Grid files
 * grid/grid_{x,y,z}_b_NNN.bin in the working directory are read at context
 * creation in place of the generated grid (load_grid, main.f90:419-428).    */
int snac_save_grid(const snac_ctx *ctx, const char *datadir);
/* initflow / initscal (initflow.f90:10-262) for every block's block_inivel
 * and every scalar's iniscal, as main.f90:531-542 calls them (is_wallturb
 * false): 'zer', 'uni', 'cou', 'poi', 'hcp', 'pdc' (laminar), 'tgv', 'kov';
 * scalars 'zer', 'uni', 'cou', 'dhc'.  'log' and 'hcl' add noise from the
 * Fortran compiler's random_number and are refused (upload those fields).
 * Ghosts and p are zero; call snac_init_bounds next.                        */
int snac_initflow(snac_ctx *ctx);
/* iparams: [iout0d, iout1d, iout2d, iout3d, isave, nsaves_max,
 *           is_overwrite_save, restart] of &dns (param.f90:88-112)         */
int snac_ctx_get_run_params(const snac_ctx *ctx, int *iparams);

/* ---- hot-path routines, all blocks at once, queued on the context's
 *      stream; results are visible after snac_sync or a download ----       */
/* rk_mom rk.f90:7-128 with momx/y/z_d, momx/y/z_a_vv, momx/y/z_p
 * (mom.f90:112-386); irk in 1..3 selects rkcoeff (param.f90:11-13).  With
 * snac_set_buoyancy on it adds the Boussinesq buoyancy of scalar 1
 * (rk.f90:104-108)                                                          */
int snac_rk_mom(snac_ctx *ctx, int irk, double dt);
/* The reference's BOUSSINESQ_BUOYANCY=1 build (rk.f90:104-108,
 * param.f90:166-170) as a run-time switch: off by default, as in a default
 * build; on, rk_mom adds -f12*gacc*beta*(face average of scalar 1).  Needs
 * nscal > 0 ("_BOUSSINESQ_BUOYANCY requires nscal > 0." otherwise).  Drops
 * the captured step graph.                                                  */
int snac_set_buoyancy(snac_ctx *ctx, int enable);
/* rk_scal rk.f90:130-204 (scal_d / scal_a scal.f90:89-155) followed by
 * bulk_mean_s / bulk_forcing_s scal.f90:157-199 for every scalar, as
 * main.f90:908-913; ghosts are filled by snac_boundp(SNAC_SCAL0 + 2*i)      */
int snac_rk_scal(snac_ctx *ctx, int irk, double dt);
/* the bulk forcing f of every scalar from the last snac_rk_scal (0 for the
 * scalars that are not forced); nscal doubles; synchronises                 */
int snac_scalar_forcing(snac_ctx *ctx, double *f);
/* bounduvw bound.f90:8-66 (updthalo 587-650 becomes device plane copies)    */
int snac_bounduvw(snac_ctx *ctx, int is_correc);
/* inflow bound.f90:438-477                                                  */
int snac_inflow(snac_ctx *ctx, int is_correc);
/* boundp bound.f90:68-106; fid = SNAC_P, SNAC_PP or a scalar
 * (SNAC_SCAL0 + 2*i, with that scalar's cbcscal/bcscal)                     */
int snac_boundp(snac_ctx *ctx, int fid);
/* fillps fillps.f90:7-54 followed by updt_rhs bound.f90:543-585             */
int snac_fillps(snac_ctx *ctx, double dtrk);
/* fft(arrplan_p(1)) + solve_n_helmholtz_2d (or solve_helmholtz) +
 * fft(arrplan_p(2)) + normfft scaling, main.f90:1004-1022                   */
int snac_solve_pressure(snac_ctx *ctx);
/* correc correc.f90:7-48                                                    */
int snac_correc(snac_ctx *ctx, double dtrk);
/* updt_pressure updt_pressure.f90:7-47 (explicit-diffusion form)            */
int snac_updt_pressure(snac_ctx *ctx);
/* chkdiv chkdiv.f90:8-44 (synchronises)                                     */
int snac_chkdiv(snac_ctx *ctx, double *divtot, double *divmax);
/* chkdiv of one block whose fields the caller holds in device memory,
 * allocating nothing.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc).  Same kernels and summation order
 * as snac_chkdiv on a one-block context (chkdiv.f90:8-44): the partial sums
 * and maxima go to a fixed device array of the library (2 x
 * SNAC_NOALLOC_MAX_PARTS doubles, 4 MiB, one copy per GPU, held until the
 * process ends) instead of context memory.  Limit: the block may need at
 * most SNAC_NOALLOC_MAX_PARTS partials, ceil(nx/64) * ceil(ny/4) * nz;
 * larger blocks are refused with an error before any launch.
 * u, v, w: ghosted fields, element (i, j, k) (0-based, ghosts at 0 and n+1)
 * at i + sx*(j + sy*k), sx >= nx+2, sy >= ny+2; dxf, dyf, dzf: the n+2
 * ghosted face spacings of each axis; out: 2 doubles of device memory that
 * receive divtot and divmax.  The launches are queued on `stream` (a
 * hipStream_t of `device`, or NULL for the null stream) and the call
 * returns without waiting.  Every call on one GPU uses that GPU's one
 * scratch array: calls that may overlap on the same GPU (other streams,
 * other host threads) race and must be ordered by the caller.             */
#define SNAC_NOALLOC_MAX_PARTS 262144
int snac_chkdiv_noalloc(int device, int nx, int ny, int nz, long sx, long sy, const double *u, const double *v,
                        const double *w, const double *dxf, const double *dyf, const double *dzf, double vol_all,
                        double *out, void *stream);
/* chkdt chkdt.f90:8-56 (synchronises); without the diffusive limit while
 * implicit diffusion is on (chkdt.f90:51-55)                                */
int snac_chkdt(snac_ctx *ctx, double *dtmax);

/* ---- implicit diffusion: the reference's _IMPDIFF build as a run-time
 *      switch.  On, snac_rk_mom keeps the diffusion out of dudtrk and half
 *      of it in the update (rk.f90:29-127), snac_step solves the Helmholtz
 *      problem (I - visc*dtrk/2 L) u = u* of each velocity component after
 *      rk_mom (main.f90:947-997) and adds alpha*L(pp) in updt_pressure
 *      (updt_pressure.f90:30-45), and snac_chkdt drops the diffusive limit.
 *      Scalars alike: rk_scal keeps half its diffusion (rk.f90:151-202)
 *      and each scalar's Helmholtz problem is solved after its bulk forcing
 *      (main.f90:914-933).  Needs the tensor-product solver (a box-shaped
 *      case with uniform transformed axes); a staggered pair along a
 *      transformed axis must be periodic or Dirichlet-Dirichlet.  ---- */
int snac_set_impdiff(snac_ctx *ctx, int enable);
/* solve_impdiff_field solve_helmholtz.f90:30-136 of fid = SNAC_U, SNAC_V,
 * SNAC_W or a scalar (SNAC_SCAL0 + 2*i) with alpha = -visc*dtrk/2 (a
 * scalar's -alpha_i*dtrk/2; < 0): phi * alphai * dx2 + BC terms
 * (init_bc_rhs solver.f90:44-157, updt_rhs bound.f90:543-585), solved in
 * place over the component's unknowns (the bounded upper face along its own
 * axis excluded)                                                            */
int snac_solve_impdiff(snac_ctx *ctx, int fid, double alpha);
/* CG iterations and largest relative residual over the modes of the last
 * MG-PCG Helmholtz solve of field fid (as snac_solver_stats for the
 * pressure: HYPRE's StructPCGGetNumIterations / GetFinalRelativeResidualNorm
 * after the Helmholtz solve of solve_helmholtz.f90:83-99); synchronises.
 * Both are 0 when that field was solved on the tensor-product plan.         */
int snac_impdiff_stats(snac_ctx *ctx, int fid, int *iters, double *max_relres);
/* updt_pressure updt_pressure.f90:7-47 of the _IMPDIFF build:
 * p += pp + alpha * Laplacian(pp), alpha = -visc*dtrk/2                     */
int snac_updt_pressure_impdiff(snac_ctx *ctx, double alpha);
/* bounduvw + boundp + inflow of the initial fields, main.f90:562-639        */
int snac_init_bounds(snac_ctx *ctx);
/* one RK3 time step in the order of main.f90:906-1033                       */
int snac_step(snac_ctx *ctx, double dt);
/* nsteps time steps with dt fixed (no host synchronisation in between)      */
int snac_steps(snac_ctx *ctx, double dt, int nsteps);
/* snac_step/snac_steps replay one hipGraph-captured step (default on; the
 * graph is re-captured when dt changes).  enable = 0 launches eagerly.
 * While phase timing is on, steps always run eagerly.                       */
int snac_set_graph(snac_ctx *ctx, int enable);
/* waits for the context's stream                                            */
int snac_sync(snac_ctx *ctx);

/* ---- per-phase timers (the reference's _TIMING hook, main.f90:1136-1143,
 *      split by phase).  While enabled, every phase of snac_step is
 *      bracketed by HIP events on the context's stream. ---- */
enum {
  SNAC_PH_RK_MOM = 0,   /* fused rk_mom                                     */
  SNAC_PH_BOUND = 1,    /* bounduvw / boundp / inflow ghost fills           */
  SNAC_PH_FILLPS = 2,   /* fillps + updt_rhs                                */
  SNAC_PH_R2R_FWD = 3,  /* forward transforms (FFT axis first)              */
  SNAC_PH_TRIDIAG = 4,  /* per-mode solves: tridiagonal lines or MG-PCG      */
  SNAC_PH_R2R_BWD = 5,  /* inverse transforms and normfft scaling           */
  SNAC_PH_CORREC = 6,   /* correc + updt_pressure                           */
  SNAC_PH_RK_SCAL = 7,  /* rk_scal + bulk forcing of the scalars            */
  SNAC_PH_HELMHOLTZ = 8, /* implicit diffusion: the three Helmholtz solves  */
  SNAC_PH_MG_LINE0 = 9, /* MG-PCG: each finest-level zebra colour sweep that
                           reads its neighbour lines (inside SNAC_PH_TRIDIAG) */
  SNAC_PH_MG_TAIL = 10, /* MG-PCG: each launch of the fused coarse end of the
                           V-cycle (inside SNAC_PH_TRIDIAG)                  */
  SNAC_PH_MG_UPDATE = 11, /* MG-PCG: each CG update pass (x += alpha p,
                           r -= alpha Mp, r.r; inside SNAC_PH_TRIDIAG)       */
  SNAC_NPHASE = 12
};
int snac_set_timing(snac_ctx *ctx, int enable);
/* waits for the stream, returns accumulated milliseconds and the number of
 * timed intervals per phase since the last call, then resets them          */
int snac_get_timing(snac_ctx *ctx, double *ms, int *count);

/* Choices of the autotuner that runs at context creation -- once per problem
 * shape in a process: a later context of the same block and transform
 * shapes reuses them, so the two run the same kernels and agree bit for bit
 * (tile shape of each
 * transform call site: forward ones, then inverse ones, then the
 * tridiagonal variant -- these only with the tensor-product solver; -1 =
 * dense path -- and last the rk_mom kernel variant: 0 one cell per thread,
 * 1..6 k-march, (v-1)%3 selecting 8/16/32 planes and (v-1)/3 the tile
 * 64x4 or 32x8; 7..10 k-march with the planes staged in LDS, 64x4 tiles
 * over 16/32 planes (7, 8) or 64x8 tiles over 16/32 planes (9, 10);
 * 11..14 as 7..10 with non-temporal hints on the read-once / write-once
 * streams; 15, 16 64x8 tiles over 64 planes without / with the hints;
 * then the FP32 rk_mom's variant of the mixed-precision modes, 7..16 as
 * above) and the best times in ms.  No
 * counterpart in the reference (its FFTW plans use FFTW_ESTIMATE).          */
int snac_get_autotune(snac_ctx *ctx, int *choices, double *ms, int n);
/* the MG-PCG line relaxation kernel each level runs (SNAC_KV_MG_LINE
 * numbering, chosen by timing when the MG-PCG plan is set up), for levels
 * 0..n-1 of the FP64 and of the FP32 V-cycle; -1 for levels in the fused
 * coarse-end kernel, past the last level, or before the MG-PCG path is
 * planned (either pointer may be null).  No counterpart in the reference. */
int snac_mg_line_choices(snac_ctx *ctx, int *fp64, int *fp32, int n);
/* overrides an autotuned kernel variant: SNAC_KV_RK_MOM (0..16, as above),
 * SNAC_KV_TRIDIAG (0: c' in LDS, lines up to 128; 1: c' in scratch; 2:
 * chunked, no scratch, lines of 8..256; 3: as 2 with twice the chunks,
 * lines of 16..128) or SNAC_KV_GHOST (0: ghost fills as
 * one launch per phase of the reference's fill order; 1, the default: each
 * fill composed at set-up into one launch that computes every ghost cell's
 * final value through the same chain of operations, where the fill allows
 * it); every variant computes the same result; SNAC_KV_RK_MOM32 (7..16):
 * the FP32 rk_mom of the mixed-precision modes; SNAC_KV_MG_LINE (process-
 * wide, every context): the MG-PCG line relaxation of the launched levels,
 * 0 each lane reading its rows from memory, 1 the rows' right-hand sides
 * and x staged through LDS with coalesced loads and stores, 2 as 1 with a
 * workgroup's lines at one position over several modes sharing the
 * coefficients it stages in LDS (FFT path; elsewhere 1), 3 and 4 as 2 with
 * 16 and 8 lanes per line instead of 32 (longer row chunks per lane, a
 * shorter interface system; lines of 80..272 and 72..264 rows, elsewhere
 * the next lower variant), a negative value (the default) each level's
 * fastest, timed when the MG-PCG plan is set up and kept per level shape
 * for the process; 0..2 compute the same values, 3 and 4 the same systems
 * in another elimination order.  No counterpart in the reference.          */
/* SNAC_KV_MG_RCP (process-wide): 1 (the default) makes the mode-batched MG
 * line sweeps divide by refined hardware reciprocals (within a few ulp of
 * the IEEE quotient); 0 divides by IEEE quotients, as the other line kernels
 * and the fused tail always do.  The V-cycle only preconditions a CG that
 * the FP64 residual stops, so both end at the same tolerance.             */
/* SNAC_KV_MG_TAIL (process-wide): 1 (the default) runs the fused coarse
 * end of the MG-PCG V-cycle with one mode's vectors of every coarse level
 * resident in LDS where they fit (160 KB per workgroup), 0 with them in
 * global memory; both compute the same values.                             */
/* SNAC_KV_MG_WARM (this context): 1 starts each MG-PCG pressure solve from
 * the previous solve's solution, as the reference's HYPRE solves start from
 * po (solver.f90:866-876, 909-920); 0 (the default) starts from zero: on
 * the channel the warm start saves no CG iteration and costs one more pass
 * over the vectors.  Both stop at the same relative residual hypre_tol.    */
/* SNAC_KV_MG_CG (process-wide): 1 (the default) forms a workgroup's new
 * CG search directions once into LDS for the operator of the direction pass
 * where the level allows it (coefficients independent of the position
 * across lines, line length dividing 256); 0 forms each neighbour's
 * direction again from z and the previous direction.  Same values.         */
enum { SNAC_KV_RK_MOM = 0, SNAC_KV_TRIDIAG = 1, SNAC_KV_GHOST = 2, SNAC_KV_RK_MOM32 = 3, SNAC_KV_MG_LINE = 4,
       SNAC_KV_MG_RCP = 5, SNAC_KV_MG_TAIL = 6, SNAC_KV_MG_WARM = 7, SNAC_KV_MG_CG = 8 };
int snac_set_kernel_variant(snac_ctx *ctx, int kernel, int variant);

/* ---- building blocks, exposed for their own parity tests ----
 * Batched r2r transform of howmany lines of length n, kinds numbered as
 * FFTW's (fftw.f90:31-54), on the device; host in/out.  strided = 0: line l
 * is contiguous at in[l*n]; 1: element q of line l is in[q*howmany + l]
 * (the layout of a transform along a slow axis).  tile = FFT tile choice,
 * as the autotuner's: 0..2 (2^-tile of the most lines per workgroup); 3, 4
 * (R2HC/HC2R of n >= 8 only, the others treat them as 0): full tiles, 2 or 4
 * per workgroup with the next tile's loads overlapping the arithmetic; 5
 * (R2HC/HC2R of n >= 8 only): full tiles, each wavefront transforming its
 * own quarter of the lines with no workgroup barrier between the passes; 6
 * (R2HC/HC2R of n >= 8 only): full tiles with the n/2-point FFT's two
 * factors in the other order; 7: as 5 with the factors in the other order. */
int snac_r2r_lines(snac_ctx *ctx, int kind, int n, int howmany, int strided, int tile, const double *in,
                   double *out);
/* STREAM-like copy of n doubles (read n*8 + write n*8 bytes per repetition)
 * on the context's stream, timed with HIP events; returns GB/s.  The HBM
 * reference of the roofline and of the byte-counter calibration; it has no
 * counterpart in the reference.                                             */
int snac_stream_copy(snac_ctx *ctx, long n, int reps, double *gbps);

/* ---- the learned preconditioner's network on its own (no counterpart in
 *      the reference: its Poisson solver is HYPRE's MG-PCG).  A per-mode
 *      convolutional encoder-decoder (4 levels, 16-64 channels, 3x3 convs,
 *      BF16 MFMA operands with FP32 accumulation, FP32 master weights,
 *      Adam).  Planes are H x W with H, W multiples of 8; inputs x are B
 *      planes of two channels ([b][y][x][2]: the scaled right-hand side and
 *      the mode's eigenvalue feature), outputs and targets B planes.  The
 *      loss is mean_b |y_b - t_b|^2 / |t_b|^2; parameters are flattened per
 *      layer as weights [tap][cout][cin] then biases [cout]. ---- */
typedef struct snac_net snac_net;
int  snac_net_create(snac_net **net, int H, int W, int maxbatch, unsigned seed, int device);
int  snac_net_destroy(snac_net *net);
long snac_net_num_params(const snac_net *net);
/* convolution kernel of every network: 0 per-tap reads from memory, 1 the
 * input staged in LDS as BF16 (default)                                   */
int  snac_net_conv_variant(int variant);
int  snac_net_get_params(const snac_net *net, float *params);
int  snac_net_set_params(snac_net *net, const float *params);
int  snac_net_forward(snac_net *net, int B, const float *x, float *y);
/* loss and its gradient w.r.t. every parameter (grad may be NULL) */
int  snac_net_loss_grad(snac_net *net, int B, const float *x, const float *t, float *loss, float *grad);
/* epochs of Adam over the samples in order, mini-batches of `batch`;
 * loss_per_epoch (may be NULL) receives each epoch's mean batch loss      */
int  snac_net_train(snac_net *net, int nsamples, const float *x, const float *t, int batch, int epochs, float lr,
                    float *loss_per_epoch);

#ifdef __cplusplus
}
#endif
#endif
