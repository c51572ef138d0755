// Batched multigrid-preconditioned conjugate gradients for the pressure
// Poisson problems on any union of blocks: the per-mode 2D Helmholtz
// problems of the FFT path (one CG per Fourier mode, all modes at once) or
// the single 3D problem without FFT.  This takes the place of HYPRE's
// per-mode StructPFMG solves (solver.f90:843-922) and of the 3D solve
// (solver.f90:524-573).
//
// Unknowns live on the bounding box of the union; cells outside it are
// inactive identity rows kept at zero.  Axes: s = the line axis (relaxed by
// line solves, never coarsened, contiguous in memory), a = the coarsened
// in-plane axis, m = the mode index (FFT path) or the third spatial axis
// (3D path, coupled and coarsened like a).  Vector element (m, a, s) sits at
// (m*na + a)*ns + s.
//
// Sign convention: the kernels work with M = -A, A being the reference's
// assembled operator (init_matrix_2d + lambda*dxf*dyf on the diagonal,
// solver.f90:687-842,981-1020; init_matrix_3d 158-372), so M is symmetric
// positive (semi-)definite:
//   (M x)(m,a,s) = (dg - lam[m]*w) x - cs(s-1) x(s-1) - cs(s) x(s+1)
//                  - ca(a-1) x(a-1) - ca(a) x(a+1) [- cm(m-1) x(m-1) - cm(m) x(m+1)]
// with c?(i) the coupling of the link i -> i+1 stored at cell i.
#pragma once
#include <functional>
#include <string>
#include <vector>

#include "case.hpp"
#include "common.hpp"

namespace snac {

class Net;

// one level, passed by value to the kernels; T is the element type of the
// level's vectors and coefficients (double, or float for the mixed-precision
// V-cycle)
template <typename T>
struct MGLevT {
  int nm, na, ns;
  int coupled;          // couplings along m (3D path)
  int per_a, per_m;     // periodic links along a / m
  long qm;              // coefficient stride of m: na*ns when coupled, else 0 (mode-independent)
  const T *dg, *w, *cs, *ca, *cm;
  const unsigned char *act;
  int nch, base, extra; // a line of ns rows over nch lanes: base (+1 for the first `extra`) rows each
  int lpl;              // lanes per line segment (32 or 64; nch <= lpl)
  int lvar;             // line relaxation kernel of this level (0..4, chosen by timing at set-up)
  // FFT path: dg, w, cs, ca and act are the same at every position a
  // (uniform, periodic in a), so a pass can read them once per row s
  int auni;
  // FFT path: the modes still iterating, in increasing order (mlist[0 ..
  // *mcount - 1]), kept by mg_flags; the mode-batched kernels take their
  // modes from it so that converged modes cost no workgroup slots (null on
  // the coupled 3D path: every mode belongs to the one group)
  const int *mlist, *mcount;
};
using MGLev = MGLevT<double>;

// the coarse end of the hierarchy handled by one fused kernel (FFT path)
constexpr int MG_TAIL_MAX = 12;
template <typename T>
struct MGTailT {
  int nl;                   // levels in the tail; the last one has na == 1
  MGLevT<T> lev[MG_TAIL_MAX];
  T *x[MG_TAIL_MAX], *b[MG_TAIL_MAX];
  // the LDS-resident form: one mode's x and b of level q at element offsets
  // ox[q] / ob[q] of the workgroup's LDS, lds_elems elements in all
  int ox[MG_TAIL_MAX], ob[MG_TAIL_MAX];
  int lds_elems;
};

// the V-cycle hierarchy in one element type: levels, the per-mode
// eigenvalues (null on the 3D path) and per-level x / b (level 0: z and r of
// the CG, or their FP32 shadows in mixed precision)
template <typename T>
struct MGHier {
  std::vector<MGLevT<T>> lev;
  const T *lam = nullptr;
  std::vector<T *> xl, bl;
  int tail0 = -1;           // first level handled by the fused tail kernel (-1: none)
  MGTailT<T> tail;
};

// ---- kernels (kernels_mg.hip), instantiated for double and float ----
// V-cycle over the tail levels, one workgroup per mode (b of the first
// tail level is given; x of every tail level is overwritten)
template <typename T>
void mg_tail(hipStream_t st, const MGTailT<T> &t, int nm, const T *lam, int pin_s, const int *singular,
             const int *done);
// zebra line relaxation of the lines (m, a) with a % 2 == col_a (and, when
// coupled, m % 2 == col_m); zero_nb: neighbouring lines read as zero;
// pin_s >= 0: in singular groups row pin_s is replaced by x = 0; xc (FFT
// path): the neighbour lines are read with the prolongation of xc (the
// next coarser level's x, cna columns) added, as mg_prolong would store it
template <typename T>
void mg_line(hipStream_t st, const MGLevT<T> &L, const T *lam, T *x, const T *b, int col_a, int col_m, int zero_nb,
             int pin_s, const int *singular, const int *done, const T *xc = nullptr, int cna = 0);
// the launched levels' line relaxation kernel: 0 reads each lane's rows
// from memory directly, 1 stages the rows' right-hand sides and the new x
// through LDS with coalesced loads and stores, 2 is 1 with a workgroup's
// lines at one position over several modes sharing the coefficients it
// stages in LDS (FFT path; elsewhere 1); 0..2 compute the same values.
// 3 and 4 are 2 with 16 and 8 lanes per line instead of the level's 32 or
// 64 (longer row chunks per lane, a shorter interface system: the same
// systems solved in another elimination order, so values differ in the
// last bits); where a level does not allow a variant, the next lower one
// runs.  The level's own choice (lvar, timed at set-up) applies unless a
// variant is forced process-wide (mg_set_line_variant 0..4; -1 returns to
// the levels' choices).  The fused tail always reads rows directly.
void mg_set_line_variant(int v);
int mg_line_variant();
// whether variant v runs as itself on level L (rather than falling back)
template <typename T>
bool mg_line_variant_runs(const MGLevT<T> &L, int v);
// times the line relaxation kernels on every launched level of a plan's
// hierarchies (FP64 and FP32) and keeps the fastest in each level's lvar
struct MGPlan;
void mg_tune_lines(MGPlan &P, hipStream_t st);
// process-wide: 1 (the default) runs the fused tail with one mode's vectors
// of every tail level resident in LDS where they fit (160 KB per
// workgroup), 0 with them in global memory; both compute the same values
void mg_set_tail_lds(bool on);
// process-wide: the CG direction pass with a workgroup's new directions
// staged in LDS where the level allows it (the default), or formed for every
// neighbour from z and p_prev; both compute the same values
void mg_set_cg_lds(bool on);
bool mg_tail_lds();
// process-wide: the mode-batched sweeps and the fused tail divide by
// refined hardware reciprocals (the estimate plus Newton steps: two in
// FP64, one in FP32) instead of IEEE divisions; off by default
void mg_set_fast_rcp(bool on);
bool mg_fast_rcp();
// coarse b = restriction of the fine residual b - M x (sum over the active
// fine cells of each coarse cell)
template <typename T>
void mg_restrict_residual(hipStream_t st, const MGLevT<T> &F, const MGLevT<T> &C, const T *lam_f, const T *x,
                          const T *b, T *bc, const int *done);
// fine x += coarse x on the active fine cells
template <typename T>
void mg_prolong(hipStream_t st, const MGLevT<T> &F, const MGLevT<T> &C, const T *xc, T *xf, const int *done);
template <typename T>
void mg_zero(hipStream_t st, long n, T *x);
// x = a fixed pseudo-random pattern in [-1, 1) (timing trials)
template <typename T>
void mg_pattern(hipStream_t st, long n, T *x);
// q = M p with out[g*2] = p.q per group, in one pass over the vectors (the
// per-group sums in mg_dot's order)
void mg_apply_dot(hipStream_t st, const MGLev &L, const double *lam, const double *p, double *q, int ng, long gsize,
                  double *part, int nchunk, double *out, const int *done);
// warm start from x0 in x with q = M x0: r = b - q (r holds b), rf (when
// given) = FP32 r and out[g*2] = r.r on the groups with b.b = bb[g*2] > 0;
// the groups with b = 0 get x = 0 (their solution), rf = 0 and r.r = 0
void mg_warm_residual(hipStream_t st, int ng, long gsize, const double *bb, double *x, double *r, const double *q,
                      float *rf, double *part, int nchunk, double *out);
// x += alpha p ; r -= alpha q with alpha = rz/pq per group, rf (when given)
// = the new r rounded to FP32, and out[g*2] = r.r, out[g*2+1] = zf.r (0
// when zf is null) of the new r, in one pass
void mg_cg_update_dot(hipStream_t st, int ng, long gsize, const double *rz, const double *pq, double *x,
                      const double *p, double *r, const double *q, float *rf, const float *zf, double *part,
                      int nchunk, double *out, const int *done);
// the CG's direction update and operator in one pass: p = z + beta p_prev
// (beta as mg_cg_dir; p stored as TP, p_prev and p distinct buffers) and
// out[g*2] = p . (M p) per group
template <typename TZ, typename TP>
void mg_dir_apply_dot(hipStream_t st, const MGLev &L, const double *lam, int ng, long gsize, const double *rzn,
                      const double *rzo, const double *cross, const TZ *z, const TP *p_prev, TP *p, double *part,
                      int nchunk, double *out, const int *done);
// mg_cg_update_dot with q = M p recomputed from p instead of read.  first:
// the solve's first step from x = 0, x not read (x = alpha p) and set to 0
// on the groups that are already done
template <typename TP>
void mg_update_apply_dot(hipStream_t st, const MGLev &L, const double *lam, int ng, long gsize, const double *rz,
                         const double *pq, double *x, const TP *p, double *r, float *rf, const float *zf,
                         double *part, int nchunk, double *out, const int *done, bool first = false);
void mg_izero(hipStream_t st, long n, int *x);
// dst = (float) src
void mg_to_float(hipStream_t st, long n, const double *src, float *dst);
// out[g*2] = x.x per group, as mg_dot computes it, and xf = (float) x in the
// same pass
void mg_dot_self_cvt(hipStream_t st, int ng, long gsize, const double *x, float *xf, double *part, int nchunk,
                     double *out);
// per-group dot products out[g*2] = x.y, out[g*2+1] = x2.y2 (or sum of x2
// when y2 is null), accumulated in FP64; groups are contiguous runs of gsize
// elements; groups with done[g] or !only[g] (when given) are skipped
template <typename TY, typename TX2 = TY>
void mg_dot(hipStream_t st, int ng, long gsize, const double *x, const TY *y, const TX2 *x2, const double *y2,
            double *part, int nchunk, double *out, const int *done, const int *only);
// p = z + beta p with beta = (rzn - cross) / rzo, cross = r_new . z_old in
// slot 1 of `cross` (flexible CG) or 0 when cross is null; beta = 0 when rzo
// is null
template <typename T>
void mg_cg_dir(hipStream_t st, int ng, long gsize, const double *rzn, const double *rzo, const double *cross,
               const T *z, double *p, const int *done);
// x -= (sum / nact) on the active cells of the singular groups
template <typename T>
void mg_sub_mean(hipStream_t st, const MGLev &L, T *x, const double *sums, const int *glist, int nlist, long gsize,
                 const double *nact,
                 const int *done);
// done[g] |= r.r <= tol2 b.b; left[0] = the groups not done; list (when
// given): those groups in increasing order.  it: the CG iterations done so
// far; the first call that finds every group done sets left[1] = it + 1
// (left[1] is cleared at the start of a solve), so a solve whose iterations
// ran ahead without a host check (captured in a graph) still tells how many
// it needed
void mg_flags(hipStream_t st, int ng, const double *rr, const double *bb, double tol2, int *done, int *left,
              int *list, int it);
// strided 3D copy (em x ea x es elements) with scaling; `tile` selects an
// LDS transpose for src stride(a) == 1 -> dst stride(s) == 1 (1) or the
// reverse (2); 0 = direct with lanes along s
void mg_copy3(hipStream_t st, const double *src, long sm, long sa, long ss, double *dst, long dm, long da, long ds,
              int em, int ea, int es, double scale, int tile);

// ---- host planner / solver (mg_solver.cpp) ----
struct MGPlan {
  bool ok = false;
  std::string why;
  int fa = -1;                 // FFT axis (0..2) or -1 (3D path)
  int ax_m = -1, ax_a = -1, ax_s = -1;
  int gmin[3] = {0, 0, 0}, G[3] = {1, 1, 1};
  int ng = 0;                  // CG groups: modes (FFT path) or 1
  long gsize = 0;
  MGHier<double> h64;          // FP64 V-cycle
  MGHier<float> h32;           // FP32 V-cycle of the mixed-precision mode
  // 0: FP64 throughout (PCG); 1: mixed -- the V-cycle (smoothing,
  // restriction, prolongation, coarse solves) in FP32, the CG vectors, the
  // operator, the dot products and the null-space removal of x, r in FP64,
  // flexible CG (Polak-Ribiere beta) for the FP32 preconditioner
  //  2: learned -- the preconditioner is the network (net.hpp) on the FP32
  //  residual, in the flexible CG of the mixed mode, falling back to the
  //  FP32 V-cycle for the rest of a solve when some mode's residual norm
  //  falls by less than the factor stall_ratio in an iteration (FFT path,
  //  eagerly launched solves)
  int prec = 0;
  Net *net = nullptr;
  int nH = 0, nW = 0;             // network planes (na, ns rounded up to 8)
  double lam_ref = 1.0, nact_plane = 1.0;
  // the field the plan solves for: -1 the pressure, 0..2 the Helmholtz
  // problem of a velocity component, 3 + i of scalar i (implicit diffusion);
  // a Helmholtz plan's solves read lam = lam_base + shift (mg_set_shift)
  int comp = -1;
  const double *lam_base = nullptr;
  double *lam64 = nullptr;
  float *lam32 = nullptr;
  // largest sum of a finest-level unknown's couplings across lines (along
  // a) over its weight, and whether the V-cycle stops at the finest level
  double cross_max = 0.0;
  bool smooth_only = false;
  // start each solve from the previous solve's solution (po, the initial
  // guess of solve_n_helmholtz_2d solver.f90:866-876) instead of zero
  bool warm = false;
  bool fallback = true;
  bool guess = false;             // also start each solve from the network's answer for b
  double stall_ratio = 0.5;
  int fallbacks = 0;              // solves that fell back to the V-cycle
  int *stall = nullptr, *stall_host = nullptr;
  double *rr_prev = nullptr;
  // training pairs recorded from the solves (network planes, FP32)
  float *pair_x = nullptr, *pair_t = nullptr, *train_loss = nullptr;
  int pair_cap = 0, pair_n = 0;
  bool collecting = false;
  double *x = nullptr, *r = nullptr, *z = nullptr, *p = nullptr, *q = nullptr;
  float *zf = nullptr, *rf = nullptr;
  // search directions of the CG iterations, ping-pong (the direction update
  // reads the previous one around each cell while it writes the new one):
  // FP64 (p and q) or FP32 (mixed and learned modes)
  double *pd[2] = {nullptr, nullptr};
  float *pf[2] = {nullptr, nullptr};
  double *part = nullptr, *bb = nullptr, *rr = nullptr, *pq = nullptr, *rz[2] = {nullptr, nullptr}, *sums = nullptr;
  double *nact = nullptr;
  int *done = nullptr, *singular = nullptr, *left = nullptr, *left_host = nullptr;
  int *mlist = nullptr;  // FFT path: the groups not done, in order (left of them)
  int *sing_list = nullptr;  // the singular groups (n_sing of them)
  int n_sing = 0;
  int nchunk = 1;
  int pin_s = -1;              // coarsest-level pinned row of singular groups
  bool any_singular = false;
  double tol = 1e-4;
  int maxiter = 50;
  int colors = 2;              // zebra colours per sweep (2, or 4 when coupled)
  int nlev = 0;
  int last_iters = 0;
  // fixed-iteration mode (graph capture): CG iteration count of the
  // captured part of a solve and the largest count an eagerly checked solve
  // needed; solves counted since the last query: all, those the host had to
  // continue past the captured count, those that stopped at maxiter above
  // the tolerance
  int fixed_iters = 0, max_seen = 0;
  int n_solves = 0, n_continued = 0, n_unconverged = 0;
  // FFT path: transform kinds and normalisation along the FFT axis, the
  // bounding-box work array and the transform plans (set up by the caller)
  int kind_fwd = 0, kind_bwd = 0;
  double norm = 1.0;
  R2RPlan fwd, bwd;
  double *ph = nullptr;
  std::vector<void *> owned;
  // optional timing marks around the finest level's colour sweeps that read
  // their neighbour lines (begin = true before the launch, false after)
  std::function<void(bool)> mark_line0;
  // the same around each launch of the fused coarse end, and of the CG
  // update pass
  std::function<void(bool)> mark_tail, mark_update;
};

// Builds the hierarchy for the case (returns with ok == false and a reason
// when the case cannot use it); device < 0 plans without allocating.  comp
// -1: the pressure's Poisson operator; 0..2 / 3 + i: the Helmholtz operator
// of velocity component comp / scalar i (implicit diffusion; its unknowns on
// the component's staggered positions, its own boundary conditions and
// transform along the FFT axis, which a component staggered along a bounded
// FFT axis takes over its n-1 interior faces)
void mg_plan(MGPlan &P, const Case &cs, int device, int comp = -1);
// Helmholtz plans: the per-mode eigenvalues the next solves use become
// lam_base + shift (shift = 1 / alpha of the substage), in both precisions.
// Where the shift outweighs the couplings across lines by far (cross_max <=
// 0.1 |shift|) the next solves precondition with the finest level's zebra
// line sweeps alone (smooth_only): the coarse levels would correct nothing
void mg_set_shift(MGPlan &P, hipStream_t st, double shift);
void mg_free(MGPlan &P);
// Host-side position of one CG solve: the iterations launched so far and
// the parity of the rz buffers.  A solve launched with a fixed iteration
// count (graph capture) leaves this state behind, so the host can resume it
// with convergence checks (mg_continue).
struct MGRun {
  int it = 0, cur = 0, pc = 0;
  bool net = false, record = false;
  bool x_unset = false;  // x is zero and left to the first update pass to write
};
// Solves M x = -rhs where rhs (in the plan's layout) has been copied into
// P.r scaled by -1; on return P.x holds the solution of A x = rhs.
// fixed == 0: the host checks the per-group convergence flags after every
// CG iteration (stops at tol or maxiter).  fixed > 0: exactly `fixed`
// iterations with no host synchronisation (converged groups skip their
// work), so the solve can be captured in a graph; the returned state then
// goes to mg_continue once the launches have run.
MGRun mg_solve(MGPlan &P, hipStream_t st, int fixed = 0);
// Resumes a fixed-count solve: synchronises, and while some group is above
// the tolerance and fewer than maxiter iterations ran, launches further
// iterations with a check after each.  Every solve thus stops at tol (or
// maxiter), whether its first iterations ran from a graph or not.
void mg_continue(MGPlan &P, hipStream_t st, MGRun &R);
// iterations of the last solve and the largest relative residual
// ||r_g|| / ||b_g|| over the groups (synchronises the stream)
void mg_stats(MGPlan &P, hipStream_t st, int &iters, double &max_relres);

}  // namespace snac
