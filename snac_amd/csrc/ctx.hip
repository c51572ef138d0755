// Device context and C-ABI (include/snac_amd.h).  Owns every block's fields
// in HBM, the ghost-fill programs, the Poisson plan, and drives one RK3 step
// in the order of main.f90:906-1033 on one HIP stream.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/snac_amd.h"
#include "net.hpp"
#include "bound_program.hpp"
#include "case.hpp"
#include "common.hpp"
#include "io.hpp"
#include "mg.hpp"

namespace snac {
namespace {

thread_local std::string g_err;

const double RKC[3][2] = {{32.0 / 60.0, 0.0}, {25.0 / 60.0, -17.0 / 60.0}, {45.0 / 60.0, -25.0 / 60.0}};

struct DevBlock {
  Geo g;
  Metrics m;
  double *met[12] = {};         // dxc dxf dyc dyf dzc dzf and their reciprocals (device)
  double *vel[2][3] = {};       // ping-pong velocity buffers
  double *p = nullptr, *pp = nullptr, *po = nullptr;
  double *dto[3] = {};
  float *dto32 = nullptr;       // the old tendencies in FP32 (mixed-precision stencils), same layout
  double *rhsp[6] = {};
  double *vin[3][3][2] = {};
  std::vector<double *> sv[2];  // scalar ping-pong buffers [parity][iscal]
  std::vector<double *> sdo;    // dsdtrko per scalar (interior)
  double *sdd = nullptr;        // a forced scalar's diffusion (implicit diffusion; interior)
  long ntot = 0, nin = 0;
  int part_off = 0, nparts = 0;
};

struct Program {
  std::vector<FaceOp *> dev;  // one device array per phase
  std::vector<int> nops, maxplane;
  CellSeg *seg = nullptr;     // the same program composed into one launch (null: not composable)
  int nseg = 0;
};

// The operator-dependent half of a tensor-product solve: the transforms of
// each transformed axis, their eigenvalues and the line coefficients.  The
// pressure's, and with implicit diffusion one per velocity component.
struct TensorOps {
  std::vector<R2RPlan> fwd, bwd;   // per transformed axis (n-1 points on a staggered DD axis)
  std::vector<double> norm;
  double *lam[3] = {};             // device eigenvalue arrays per axis (line axis: zeros)
  double *am = nullptr, *ap = nullptr, *d0 = nullptr, *wt = nullptr;
  int pin = 0;
};

// Tensor-product Poisson plan: r2r along every uniform axis (the FFT axis
// first), tridiagonal solves along the remaining axis.
struct TensorPlan {
  bool ok = false;
  std::string why;
  int G[3] = {1, 1, 1}, gmin[3] = {0, 0, 0};
  std::vector<int> xaxes;          // transformed axes, FFT axis first
  int s = -1;
  bool all_tr = false;             // every axis periodic: the line axis is transformed too
  std::vector<double> gdf[3], gdcm[3], gdcp[3];  // global 1D grids: df, dc below / above each cell
  TensorOps P;                     // the pressure operator
  int tvariant = 1;                // tridiagonal kernel variant (autotuned)
  double *ph = nullptr, *scratch = nullptr, *scratch2 = nullptr;
  std::vector<double> tune_ms;     // best time per tuned stage (diagnostics)
};

// Implicit diffusion of one velocity component: its tensor operator and,
// per block, the right-hand-side description (weights, BC planes)
struct ImpField {
  TensorOps ops;              // the tensor-product path
  std::vector<ImpRhs> rhs;
  std::shared_ptr<MGPlan> mg;  // the MG-PCG path: the field's own plan, built on first need
  bool mg_ok = false;          // the field's operator assembles on the MG-PCG path
};

}  // namespace
}  // namespace snac

struct snac_ctx {
  snac::Case cs;
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<snac::DevBlock> blk;
  int rk_variant = 0;     // rk_mom kernel variant (autotuned)
  double rk_tune_ms = 0.0;
  int rk32_variant = 16;  // the FP32 rk_mom's (LDS variants 7..16, autotuned)
  double rk32_tune_ms = 0.0;
  bool st32 = false;      // momentum right-hand side in FP32 (precision MIXED / LEARNED)
  int cur = 0;   // parity of the current velocity buffers
  int scur = 0;  // parity of the current scalar buffers (flips once per substage)
  snac::Program prog_uvw[2][2], prog_bound_only[2][2], prog_inflow[2][2], prog_p, prog_pp, prog_end[2];
  snac::Program prog_scal[2];                 // boundp of every scalar, per parity
  std::vector<snac::Program> prog_scal_one[2];  // boundp of one scalar
  double *sred = nullptr, *sforce = nullptr;  // scalar bulk mean / forcing
  snac::TensorPlan tp;
  bool impdiff = false;    // implicit diffusion (the reference's _IMPDIFF build)
  bool buoyancy = false;   // Boussinesq buoyancy (the reference's _BOUSSINESQ_BUOYANCY build)
  std::vector<snac::ImpField> imp;  // its Helmholtz operators of u, v, w and the scalars
  snac::MGPlan mg;
  bool mg_planned = false;
  bool mg_warm = false;  // pressure solves start from the previous solution (SNAC_KV_MG_WARM)
  int solver = SNAC_SOLVER_AUTO;
  int prec = SNAC_PREC_FP64;
  bool ghost_fused = true;  // run composed ghost-fill programs in one launch
  double *part_sum = nullptr, *part_max = nullptr, *red = nullptr;
  int nparts = 0;
  double dlmin = 0;
  std::vector<void *> owned;
  // phase timers
  bool timing = false;
  struct TRec { int ph; hipEvent_t a, b; };
  std::vector<TRec> pending;
  std::vector<hipEvent_t> pool;
  double tacc[SNAC_NPHASE] = {};
  int tcnt[SNAC_NPHASE] = {};
  // one captured RK3 step per starting scalar parity, replayed by
  // snac_step / snac_steps.  On the MG-PCG path the step is captured as
  // segments that each end after the captured iterations of a pressure
  // solve; between segments the host resumes that solve (gres: its CG state
  // at the end of the segment) until it meets the tolerance
  bool use_graph = true;
  std::vector<hipGraphExec_t> gseg[2];
  std::vector<std::pair<snac::MGPlan *, snac::MGRun>> gres[2];
  double graph_dt[2] = {0.0, 0.0};
  int graph_cur[2] = {-1, -1};
  bool capturing = false;  // a step is being captured (segments go to cap_key)
  int cap_key = 0;
  int force_fixed = 0;     // CG iterations captured per solve (0: measured)
};

namespace snac {
namespace {

// device allocations are skipped in a host-only context (device < 0)
template <class T>
T *dalloc(snac_ctx *c, size_t n) {
  if (c->device < 0) return nullptr;
  void *p = nullptr;
  SNAC_HIP(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
  SNAC_HIP(hipMemset(p, 0, std::max<size_t>(n, 1) * sizeof(T)));
  c->owned.push_back(p);
  return (T *)p;
}
// Ghosted field with the leading dimension padded to a multiple of 16
// doubles and the base shifted so that interior element (1, j, k) of every
// row starts on a 128-byte boundary (hipMalloc returns 256-byte aligned
// memory): rows of 16 or 64 neighbouring lines are whole cache lines.
double *dalloc_field(snac_ctx *c, long ntot) {
  double *p = dalloc<double>(c, (size_t)ntot + 16);
  return p ? p + 15 : nullptr;
}

template <class T>
T *dupload(snac_ctx *c, const std::vector<T> &h) {
  T *p = dalloc<T>(c, h.size());
  if (p && !h.empty()) SNAC_HIP(hipMemcpy(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  return p;
}

std::vector<BlockArrays> block_arrays(snac_ctx *c) {
  std::vector<BlockArrays> out(c->blk.size());
  for (size_t b = 0; b < c->blk.size(); b++) {
    DevBlock &D = c->blk[b];
    BlockArrays &A = out[b];
    A.g = D.g;
    for (int par = 0; par < 2; par++)
      for (int q = 0; q < 3; q++) A.vel[par][q] = D.vel[par][q];
    A.p = D.p;
    A.pp = D.pp;
    for (int par = 0; par < 2; par++) A.sv[par] = D.sv[par];
    for (int d = 0; d < 3; d++)
      for (int q = 0; q < 3; q++)
        for (int s = 0; s < 2; s++) A.vin[d][q][s] = D.vin[d][q][s];
  }
  return out;
}

Program upload_program(snac_ctx *c, const std::vector<std::vector<FaceOp>> &phases) {
  Program pr;
  if (c->device < 0) return pr;
  for (auto &ph : phases) {
    FaceOp *d = dalloc<FaceOp>(c, ph.size());
    SNAC_HIP(hipMemcpy(d, ph.data(), ph.size() * sizeof(FaceOp), hipMemcpyHostToDevice));
    int mp = 0;
    for (auto &o : ph) mp = std::max(mp, o.m1 * o.m2);
    pr.dev.push_back(d);
    pr.nops.push_back((int)ph.size());
    pr.maxplane.push_back(mp);
  }
  std::vector<CellSeg> segs;
  if (phases.size() > 1 && compose_program(phases, segs) && !segs.empty()) {
    pr.seg = dalloc<CellSeg>(c, segs.size());
    SNAC_HIP(hipMemcpy(pr.seg, segs.data(), segs.size() * sizeof(CellSeg), hipMemcpyHostToDevice));
    pr.nseg = (int)segs.size();
  }
  return pr;
}

void run_program(snac_ctx *c, const Program &pr) {
  if (pr.nseg > 0 && c->ghost_fused) {
    launch_cell_segs(c->stream, pr.seg, pr.nseg);
    return;
  }
  for (size_t q = 0; q < pr.dev.size(); q++) launch_face_ops(c->stream, pr.dev[q], pr.nops[q], pr.maxplane[q]);
}

hipEvent_t ev_get(snac_ctx *c) {
  if (!c->pool.empty()) {
    hipEvent_t e = c->pool.back();
    c->pool.pop_back();
    return e;
  }
  hipEvent_t e;
  SNAC_HIP(hipEventCreate(&e));
  return e;
}

// brackets one phase with events when timing is on
struct Phase {
  snac_ctx *c;
  int ph;
  hipEvent_t a{}, b{};
  Phase(snac_ctx *cc, int p) : c(cc), ph(p) {
    if (!c->timing) return;
    a = ev_get(c);
    SNAC_HIP(hipEventRecord(a, c->stream));
  }
  ~Phase() {
    if (!c->timing) return;
    b = ev_get(c);
    if (hipEventRecord(b, c->stream) == hipSuccess) c->pending.push_back({ph, a, b});
  }
};

void collect_timing(snac_ctx *c) {
  if (c->pending.empty()) return;
  SNAC_HIP(hipStreamSynchronize(c->stream));
  for (auto &r : c->pending) {
    float ms = 0.f;
    SNAC_HIP(hipEventElapsedTime(&ms, r.a, r.b));
    c->tacc[r.ph] += ms;
    c->tcnt[r.ph] += 1;
    c->pool.push_back(r.a);
    c->pool.push_back(r.b);
  }
  c->pending.clear();
}

// BC character of field comp (-1 pressure, 0..2 velocity component, 3+i
// scalar i) on the global face (axis a, side) of the plan's box: 'F' where
// periodic
bool box_bc(snac_ctx *c, int comp, int a, int side, char &out, std::string &why) {
  const TensorPlan &t = c->tp;
  out = 0;
  for (auto &B : c->cs.blocks) {
    const bool on = side == 0 ? B.lo_g[a] == t.gmin[a] : B.lo_g[a] + B.n[a] == t.gmin[a] + t.G[a];
    if (!on) continue;
    const char bc = !B.is_bound[a][side] ? 'F'
                    : comp < 0 ? B.cbcpre[a][side]
                    : comp < 3 ? B.cbcvel[comp][a][side]
                               : B.cbcsca[((comp - 3) * 3 + a) * 2 + side];
    if (out && out != bc) { why = "mixed boundary conditions on one face"; return false; }
    out = bc;
  }
  return true;
}

// Transforms, eigenvalues and line coefficients of one operator on the
// plan's axes: the pressure (comp -1, centred), the Helmholtz operator of a
// scalar (comp 3+i, centred) or of velocity component comp (implicit
// diffusion): staggered along its own axis
// (is_centered = .false., main.f90:774-776), i.e. unknowns on the faces (the
// bounded upper face excluded: an identity row of the line solve, or a
// transform over n-1 points), couplings 1/dxf, weights dxc, Dirichlet values
// on the boundary face itself (sgn 0).  init_matrix_2d/3d solver.f90:158-372,
// 687-842, init_fft_reduction 1124-1138, fft.f90:96-189.
bool build_ops(snac_ctx *c, TensorOps &o, int comp, std::string &why) {
  TensorPlan &t = c->tp;
  const Case &cs = c->cs;
  const int fa = cs.fft_axis - 1, s = t.s;
  char lo[3], hi[3];
  for (int a = 0; a < 3; a++)
    if (!box_bc(c, comp, a, 0, lo[a], why) || !box_bc(c, comp, a, 1, hi[a], why)) return false;
  std::vector<double> lamh[3];
  for (int a = 0; a < 3; a++) lamh[a].assign(t.G[a], 0.0);
  for (int a : t.xaxes) {
    int kf, kb;
    double nr[2];
    const bool centred = a != comp;
    char c0 = (a == fa && comp < 0) ? cs.blocks[0].cbcpre[a][0] : lo[a];
    char c1 = (a == fa && comp < 0) ? cs.blocks[0].cbcpre[a][1] : hi[a];
    if (!find_fft(c0, c1, kf, kb, nr, centred)) {
      why = std::string("unsupported BC pair ") + c0 + c1 + (centred ? "" : " (staggered)") +
            " along a transformed axis";
      return false;
    }
    const int q = (!centred && c0 != 'F') ? 1 : 0;  // fftini fft.f90:31-33
    if (c->device >= 0) {
      o.fwd.push_back(r2r_plan(kf, t.G[a] - q));
      o.bwd.push_back(r2r_plan(kb, t.G[a] - q));
    }
    o.norm.push_back(1.0 / (nr[0] * (t.G[a] + nr[1] - q)));
    eigenvalues(t.G[a], c0, c1, lamh[a], centred);
    // spacing used by the reference: dxc_g(lo_g-1) of the FFT axis
    // (main.f90:717-728, dxf_g(lo_g) when staggered); the uniform spacing
    // for the other axes
    double h = (a == fa) ? cs.blocks[0].df[a][1] : t.gdf[a][0];
    for (auto &l : lamh[a]) l /= h * h;
  }
  for (int a = 0; a < 3; a++) o.lam[a] = dupload(c, lamh[a]);
  // line coefficients: h_other = product of the uniform spacings of the
  // transformed in-plane axes (the FFT axis is not part of the 2D operator)
  double hother = 1.0;
  for (int a = 0; a < 3; a++) if (a != fa && a != s) hother *= t.gdf[a][0];
  const bool stag = comp == s;
  const int m = t.G[s];
  const int mu = (stag && hi[s] != 'F') ? m - 1 : m;  // unknowns along the line
  std::vector<double> am(m), ap(m), d0(m), wt(m);
  auto sgn = [&](char bc) { return bc == 'N' ? 1.0 : (bc == 'D' ? (stag ? 0.0 : -1.0) : 0.0); };
  const double sg0 = sgn(lo[s]), sg1 = sgn(hi[s]);
  for (int g = 0; g < m; g++) {
    wt[g] = hother * (stag ? t.gdcp[s][g] : t.gdf[s][g]);
    if (t.all_tr) {  // the line axis is transformed as well: a diagonal system
      am[g] = ap[g] = 0.0;
      d0[g] = lamh[s][g] * wt[g];
      continue;
    }
    if (g >= mu) {  // the bounded upper face of a staggered component
      am[g] = ap[g] = 0.0;
      d0[g] = 1.0;
      wt[g] = 0.0;
      continue;
    }
    const double rm = 1.0 / (stag ? t.gdf[s][g] : t.gdcm[s][g]) * hother;
    const double rp = 1.0 / (stag ? t.gdf[s][g + 1] : t.gdcp[s][g]) * hother;
    double dg = -(rm + rp);
    am[g] = rm; ap[g] = rp;
    if (g == 0) { dg += sg0 * rm; am[g] = 0.0; }
    if (g == mu - 1) { dg += sg1 * rp; ap[g] = 0.0; }
    d0[g] = dg;
  }
  o.am = dupload(c, am); o.ap = dupload(c, ap); o.d0 = dupload(c, d0); o.wt = dupload(c, wt);
  // singular N-N line of the pressure: pin its last unknown; the
  // all-periodic zero mode has a zero pivot instead, which the line solver
  // maps to x = 0
  o.pin = (comp < 0 && !t.all_tr && lo[s] == 'N' && hi[s] == 'N') ? 1 : 0;
  return true;
}

// ---- tensor Poisson planner ----
void plan_tensor(snac_ctx *c) {
  TensorPlan &t = c->tp;
  const Case &cs = c->cs;
  int fa = cs.fft_axis - 1;
  // global extents of the union
  int gmax[3];
  for (int a = 0; a < 3; a++) {
    t.gmin[a] = 1 << 30;
    gmax[a] = -(1 << 30);
    for (auto &B : cs.blocks) {
      t.gmin[a] = std::min(t.gmin[a], B.lo_g[a]);
      gmax[a] = std::max(gmax[a], B.lo_g[a] + B.n[a] - 1);
    }
    t.G[a] = gmax[a] - t.gmin[a] + 1;
  }
  double vol = 0;
  for (auto &B : cs.blocks) vol += (double)B.n[0] * B.n[1] * B.n[2];
  if (vol != (double)t.G[0] * t.G[1] * t.G[2]) { t.why = "the blocks do not tile a box"; return; }
  // per-axis global 1D grids: df[g], dcm[g] (between g-1 and g), dcp[g]
  std::vector<double> gdf[3], gdcm[3], gdcp[3];
  char bclo[3], bchi[3];
  for (int a = 0; a < 3; a++) {
    gdf[a].assign(t.G[a], 0.0); gdcm[a].assign(t.G[a], 0.0); gdcp[a].assign(t.G[a], 0.0);
    std::vector<int> seen(t.G[a], 0);
    bclo[a] = bchi[a] = 0;
    for (auto &B : cs.blocks) {
      for (int i = 1; i <= B.n[a]; i++) {
        int g = B.lo_g[a] - t.gmin[a] + i - 1;
        double f = B.df[a][i], cm = B.dc[a][i - 1], cp = B.dc[a][i];
        if (seen[g] && (std::fabs(gdf[a][g] - f) > 1e-14 * f || std::fabs(gdcm[a][g] - cm) > 1e-14 * cm ||
                        std::fabs(gdcp[a][g] - cp) > 1e-14 * cp)) {
          t.why = "inconsistent grids among blocks";
          return;
        }
        gdf[a][g] = f; gdcm[a][g] = cm; gdcp[a][g] = cp; seen[g] = 1;
      }
      if (B.lo_g[a] == t.gmin[a]) {
        char bc = B.is_bound[a][0] ? B.cbcpre[a][0] : 'F';
        if (bclo[a] && bclo[a] != bc) { t.why = "mixed boundary conditions on one face"; return; }
        bclo[a] = bc;
      }
      if (B.lo_g[a] + B.n[a] - 1 == gmax[a]) {
        char bc = B.is_bound[a][1] ? B.cbcpre[a][1] : 'F';
        if (bchi[a] && bchi[a] != bc) { t.why = "mixed boundary conditions on one face"; return; }
        bchi[a] = bc;
      }
    }
    if ((bclo[a] == 'F') != (bchi[a] == 'F')) { t.why = "half-periodic axis"; return; }
  }
  // candidate transform axes: uniform spacing and a transformable BC pair
  std::vector<int> red;
  for (int a = 0; a < 3; a++) if (a != fa) red.push_back(a);
  auto uniform = [&](int a) {
    double h = gdf[a][0];
    for (int g = 0; g < t.G[a]; g++)
      if (std::fabs(gdf[a][g] - h) > 1e-12 * h || std::fabs(gdcm[a][g] - h) > 1e-12 * h ||
          std::fabs(gdcp[a][g] - h) > 1e-12 * h)
        return false;
    return true;
  };
  auto transformable = [&](int a) {
    int kf, kb;
    double nr[2];
    return find_fft(bclo[a], bchi[a], kf, kb, nr);
  };
  std::vector<int> cand, non;
  for (int a : red) ((uniform(a) && transformable(a)) ? cand : non).push_back(a);
  if (non.size() > 1) { t.why = "more than one non-uniform axis"; return; }
  int s;
  bool all_tr = false;  // every axis periodic: transformed too, diagonal "lines"
  if (non.size() == 1) s = non[0];
  else {
    s = -1;
    for (int q = (int)cand.size() - 1; q >= 0; q--)
      if (bclo[cand[q]] != 'F') { s = cand[q]; break; }
    if (s < 0) {
      if (cand.empty()) { t.why = "no axis for the line solves"; return; }
      s = cand.back();
      all_tr = true;
    }
  }
  if (bclo[s] == 'F' && !all_tr) { t.why = "periodic line axis"; return; }
  t.s = s;
  if (fa >= 0) t.xaxes.push_back(fa);
  for (int a : red) if (a != s) t.xaxes.push_back(a);
  if (all_tr) t.xaxes.push_back(s);
  // the first transform runs block by block from pp into the work array, so
  // its axis must be spanned by every block (always true for the FFT axis)
  {
    int first = -1;
    for (size_t q = 0; q < t.xaxes.size() && first < 0; q++) {
      bool full = true;
      for (auto &B : cs.blocks) full = full && B.n[t.xaxes[q]] == t.G[t.xaxes[q]];
      if (full) first = (int)q;
    }
    if (first < 0) { t.why = "no transform axis spans every block"; return; }
    std::swap(t.xaxes[0], t.xaxes[first]);
  }
  t.all_tr = all_tr;
  for (int a = 0; a < 3; a++) { t.gdf[a] = gdf[a]; t.gdcm[a] = gdcm[a]; t.gdcp[a] = gdcp[a]; }
  if (!build_ops(c, t.P, -1, t.why)) return;
  size_t tot = (size_t)t.G[0] * t.G[1] * t.G[2];
  t.ph = dalloc<double>(c, tot);
  t.scratch = dalloc<double>(c, tot);
  t.scratch2 = dalloc<double>(c, tot);
  t.ok = true;
}

// batched MG-PCG plan with the FFT-axis transforms and the bounding-box
// work array of the FFT path; planned once, on first need
void finish_mg_plan(snac_ctx *c, MGPlan &P);

void plan_mg(snac_ctx *c) {
  if (c->mg_planned) return;
  c->mg_planned = true;
  MGPlan &P = c->mg;
  mg_plan(P, c->cs, c->device);
  P.prec = c->prec;
  P.warm = c->mg_warm;
  finish_mg_plan(c, P);
}

// after mg_plan: each level's fastest line kernel, the FFT-axis transforms
// and the bounding-box work array (a Helmholtz plan shares the pressure
// plan's: the solves run one after the other on the context's stream)
void finish_mg_plan(snac_ctx *c, MGPlan &P) {
  if (P.ok && c->device >= 0) mg_tune_lines(P, c->stream);
  if (!P.ok || c->device < 0 || P.fa < 0) return;
  // ng modes along the FFT axis (its n-1 first planes for a component
  // staggered along a bounded FFT axis)
  P.fwd = r2r_plan(P.kind_fwd, P.ng);
  P.bwd = r2r_plan(P.kind_bwd, P.ng);
  if (P.comp >= 0) {
    plan_mg(c);
    if (c->mg.ok && c->mg.ph) {
      P.ph = c->mg.ph;
      return;
    }
  }
  P.ph = dalloc<double>(c, (size_t)P.G[0] * P.G[1] * P.G[2]);
}

// true when the pressure solve runs the MG-PCG path
bool use_mg(snac_ctx *c) {
  if (c->solver == SNAC_SOLVER_MG) return true;
  if (c->solver == SNAC_SOLVER_TENSOR) return false;
  return !c->tp.ok;
}

void check_solver(snac_ctx *c) {
  if (c->solver == SNAC_SOLVER_TENSOR && !c->tp.ok)
    throw std::runtime_error("the tensor-product solver needs a box-shaped union with uniform transform axes (" +
                             c->tp.why + ")");
  if (use_mg(c)) {
    plan_mg(c);
    if (!c->mg.ok)
      throw std::runtime_error("no GPU Poisson solver for this case: tensor-product solver: " + c->tp.why +
                               "; multigrid solver: " + c->mg.why);
  }
}

// strides of axis a in a ghosted block array, and of the interior origin
long gstride(const DevBlock &D, int a) { return a == 0 ? 1 : (a == 1 ? D.g.sx : D.g.sx * D.g.sy); }
int bext(const DevBlock &D, int a) { return a == 0 ? D.g.nx : (a == 1 ? D.g.ny : D.g.nz); }

void others(int a, int &b1, int &b2) {
  b1 = a == 0 ? 1 : 0;
  b2 = a == 2 ? 1 : 2;
}

// first transform: block pp interiors -> global work array
void stage_first(snac_ctx *c, const TensorOps &o) {
  TensorPlan &t = c->tp;
  const long pst[3] = {1, (long)t.G[0], (long)t.G[0] * t.G[1]};
  int a = t.xaxes[0], b1, b2;
  others(a, b1, b2);
  for (size_t b = 0; b < c->blk.size(); b++) {
    DevBlock &D = c->blk[b];
    const Block &B = c->cs.blocks[b];
    const double *in = D.pp + 1 + D.g.sx + D.g.sx * D.g.sy;
    long off = 0;
    for (int q = 0; q < 3; q++) off += (long)(B.lo_g[q] - t.gmin[q]) * pst[q];
    launch_r2r(c->stream, o.fwd[0], in, gstride(D, a), gstride(D, b1), gstride(D, b2), t.ph + off, pst[a], pst[b1],
               pst[b2], bext(D, b1), bext(D, b2), 1.0);
  }
}

// in-place transform number q on the work array (forward or inverse)
void stage_inplace(snac_ctx *c, const TensorOps &o, size_t q, bool fwd) {
  TensorPlan &t = c->tp;
  const long pst[3] = {1, (long)t.G[0], (long)t.G[0] * t.G[1]};
  int a = t.xaxes[q], b1, b2;
  others(a, b1, b2);
  launch_r2r(c->stream, fwd ? o.fwd[q] : o.bwd[q], t.ph, pst[a], pst[b1], pst[b2], t.ph, pst[a], pst[b1], pst[b2],
             t.G[b1], t.G[b2], fwd ? 1.0 : o.norm[q]);
}

void stage_tridiag(snac_ctx *c, const TensorOps &o, double shift) {
  TensorPlan &t = c->tp;
  TridiagLine tl;
  for (int q = 0; q < 3; q++) tl.G[q] = t.G[q];
  tl.s = t.s;
  tl.variant = t.tvariant;
  int o1 = t.s == 0 ? 1 : 0, o2 = t.s == 2 ? 1 : 2;
  tl.am = o.am; tl.ap = o.ap; tl.d0 = o.d0; tl.wt = o.wt;
  tl.lamA = o.lam[o1];
  tl.lamB = o.lam[o2];
  tl.pin_when_zero = o.pin;
  tl.shift = shift;
  launch_tridiag(c->stream, tl, t.ph, t.scratch, t.scratch2);
}

// the current buffer of field comp of a Helmholtz solve (velocity
// component 0..2, scalar 3+i)
double *imp_field(snac_ctx *c, DevBlock &D, int comp) {
  return comp < 3 ? D.vel[c->cur][comp] : D.sv[c->scur][comp - 3];
}

// last inverse transform: global work array -> block interiors (pp for the
// pressure, field comp for a Helmholtz solve), scaled
void stage_last(snac_ctx *c, const TensorOps &o, int comp) {
  TensorPlan &t = c->tp;
  const long pst[3] = {1, (long)t.G[0], (long)t.G[0] * t.G[1]};
  int a = t.xaxes[0], b1, b2;
  others(a, b1, b2);
  for (size_t b = 0; b < c->blk.size(); b++) {
    DevBlock &D = c->blk[b];
    const Block &B = c->cs.blocks[b];
    double *out = (comp < 0 ? D.pp : imp_field(c, D, comp)) + 1 + D.g.sx + D.g.sx * D.g.sy;
    long off = 0;
    for (int q = 0; q < 3; q++) off += (long)(B.lo_g[q] - t.gmin[q]) * pst[q];
    launch_r2r(c->stream, o.bwd[0], t.ph + off, pst[a], pst[b1], pst[b2], out, gstride(D, a), gstride(D, b1),
               gstride(D, b2), bext(D, b1), bext(D, b2), o.norm[0]);
  }
}

// the whole tensor-product solve of one operator, pp -> pp (pressure) or
// pp -> velocity component comp
void tensor_solve(snac_ctx *c, const TensorOps &o, int comp, double shift) {
  TensorPlan &t = c->tp;
  {
    Phase tf(c, comp < 0 ? SNAC_PH_R2R_FWD : SNAC_PH_HELMHOLTZ);
    stage_first(c, o);
    for (size_t q = 1; q < t.xaxes.size(); q++) stage_inplace(c, o, q, true);
  }
  {
    Phase tt(c, comp < 0 ? SNAC_PH_TRIDIAG : SNAC_PH_HELMHOLTZ);
    stage_tridiag(c, o, shift);
  }
  {
    Phase tb(c, comp < 0 ? SNAC_PH_R2R_BWD : SNAC_PH_HELMHOLTZ);
    for (size_t q = t.xaxes.size() - 1; q >= 1; q--) stage_inplace(c, o, q, false);
    stage_last(c, o, comp);
  }
}

// Implicit diffusion on the tensor-product plan (the reference's _IMPDIFF
// build, main.f90:769-886): per velocity component its tensor operator and,
// per block, the weights dx2 of the non-FFT axes and the BC planes of
// init_bc_rhs (solver.f90:44-157) over the component's unknowns: factor *
// weight / (dx1(i-(1-ib)+qqq) * dx2(i)), with factor dl*bc (Neumann, negated
// on the upper face), -2 bc (centred Dirichlet) or -bc (staggered
// Dirichlet), times inflow / bc where an inflow plane gives the boundary
// values (main.f90:600-636).
void plan_impdiff(snac_ctx *c) {
  TensorPlan &t = c->tp;
  const Case &cs = c->cs;
  const int fa = cs.fft_axis - 1;
  c->imp.assign(3 + cs.nscal, ImpField());
  for (int comp = 0; comp < 3 + cs.nscal; comp++) {
    ImpField &f = c->imp[comp];
    const std::string name = comp < 3 ? std::string(1, "uvw"[comp]) : "scalar " + std::to_string(comp - 2);
    std::string why;
    if (t.ok) {
      if (!build_ops(c, f.ops, comp, why)) throw std::runtime_error("implicit diffusion of " + name + ": " + why);
    }
    {
      // the MG-PCG path: whether the field's operator assembles (host-side
      // check; required when there is no tensor-product plan)
      MGPlan probe;
      mg_plan(probe, cs, -1, comp);
      f.mg_ok = probe.ok;
      if (!probe.ok && !t.ok) throw std::runtime_error("implicit diffusion of " + name + ": " + probe.why);
    }
    // the tuned transform tiles of the pressure's plans where the kernels match
    for (size_t q = 0; q < f.ops.fwd.size() && q < t.P.fwd.size(); q++) {
      if (f.ops.fwd[q].kind == t.P.fwd[q].kind && f.ops.fwd[q].n == t.P.fwd[q].n) f.ops.fwd[q].lshift = t.P.fwd[q].lshift;
      if (f.ops.bwd[q].kind == t.P.bwd[q].kind && f.ops.bwd[q].n == t.P.bwd[q].n) f.ops.bwd[q].lshift = t.P.bwd[q].lshift;
    }
    for (size_t b = 0; b < cs.blocks.size(); b++) {
      const Block &B = cs.blocks[b];
      ImpRhs r{};
      int ext[3];
      for (int a = 0; a < 3; a++) ext[a] = r.ext[a] = B.n[a] - ((a == comp && B.is_bound[a][1]) ? 1 : 0);
      auto dx2 = [&](int a, int i) { return a == comp ? B.dc[a][i] : B.df[a][i]; };
      auto sp_lo = [&](int a, int i) { return a == comp ? B.df[a][i] : B.dc[a][i - 1]; };
      auto sp_hi = [&](int a, int i) { return a == comp ? B.df[a][i + 1] : B.dc[a][i]; };
      for (int a = 0; a < 3; a++) {
        std::vector<double> w(B.n[a], 1.0);
        if (a != fa)
          for (int i = 1; i <= B.n[a]; i++) w[i - 1] = dx2(a, i);
        r.w[a] = dupload(c, w);
      }
      for (int d = 0; d < 3; d++)
        for (int sd = 0; sd < 2; sd++) {
          r.plane[d][sd] = nullptr;
          if (!B.is_bound[d][sd] || cs.periods[d] != 0) continue;
          const bool vel = comp < 3;
          const size_t qs = vel ? 0 : ((size_t)(comp - 3) * 3 + d) * 2 + sd;
          const char ct = vel ? B.cbcvel[comp][d][sd] : B.cbcsca[qs];
          const double bcv = vel ? B.bcvel[comp][d][sd] : B.bcsca[qs];
          const double dl = sd == 0 ? B.df[d][1] : B.df[d][B.n[d]];
          double factor = 0.0;
          if (ct == 'N') factor = sd == 0 ? dl * bcv : -(dl * bcv);
          else if (ct == 'D') factor = d == comp ? -1.0 * bcv : -2.0 * bcv;
          const int a1 = d == 0 ? 1 : 0, a2 = d == 2 ? 1 : 2;
          const int q = sd == 0 ? 1 : ext[d];
          static const std::vector<double> none;
          const std::vector<double> &vin = vel ? B.vin[d][comp][sd] : none;
          const bool arr = vel && B.is_inflow[d][sd] && !vin.empty() && bcv != 0.0;
          std::vector<double> pl((size_t)ext[a1] * ext[a2]);
          bool any = false;
          for (int t2 = 1; t2 <= ext[a2]; t2++)
            for (int t1 = 1; t1 <= ext[a1]; t1++) {
              int ijk[3];
              ijk[d] = q; ijk[a1] = t1; ijk[a2] = t2;
              double weight = 1.0;
              for (int a = 0; a < 3; a++) if (a != fa) weight *= dx2(a, ijk[a]);
              const double spc = sd == 0 ? sp_lo(d, q) : sp_hi(d, q);
              double rhs = factor * weight / (spc * dx2(d, q));
              if (arr) rhs = rhs * vin[t1 + (size_t)(B.n[a1] + 2) * t2] / bcv;
              pl[(t1 - 1) + (size_t)ext[a1] * (t2 - 1)] = rhs;
              any = any || rhs != 0.0;
            }
          if (any) r.plane[d][sd] = dupload(c, pl);
        }
      f.rhs.push_back(r);
    }
  }
  // a forced scalar keeps its diffusion until its bulk mean is taken
  bool forced = false;
  for (int is = 0; is < cs.nscal; is++) forced = forced || cs.is_sforced[is];
  if (forced)
    for (auto &D : c->blk)
      if (!D.sdd) D.sdd = dalloc<double>(c, D.nin);
}

void mg_stage_in(snac_ctx *c, MGPlan &P);
void mg_stage_out(snac_ctx *c, MGPlan &P);
void finish_mg_plan(snac_ctx *c, MGPlan &P);
void split_capture(snac_ctx *c, MGPlan &P, const MGRun &R);

// solve_impdiff_field (solve_helmholtz.f90:30-136) of velocity component
// comp (0..2) or scalar comp-3, in place on its current buffer: the
// right-hand side goes to pp, the solve back into the field's unknowns.
// With the pressure on the MG-PCG path the field's Helmholtz problem is
// solved there too, by its own plan: the per-mode eigenvalues shifted by
// 1/alpha (the weighted diagonal term of add_weighted_constant_to_diagonal)
void solve_impdiff(snac_ctx *c, int comp, double alpha) {
  const double alphai = 1.0 / alpha;
  ImpField &f = c->imp[comp];
  {
    Phase tp(c, SNAC_PH_HELMHOLTZ);
    for (size_t b = 0; b < c->blk.size(); b++) {
      DevBlock &D = c->blk[b];
      launch_imp_rhs(c->stream, D.g, f.rhs[b], imp_field(c, D, comp), D.pp, alphai);
    }
  }
  if (!use_mg(c) || (!f.mg_ok && c->tp.ok)) {  // the tensor-product path where the field allows it
    tensor_solve(c, f.ops, comp, alphai);
    return;
  }
  if (!f.mg) {
    if (c->capturing) throw std::runtime_error("internal: a Helmholtz MG plan created during graph capture");
    auto P = std::shared_ptr<MGPlan>(new MGPlan(), [](MGPlan *q) {
      mg_free(*q);
      delete q;
    });
    mg_plan(*P, c->cs, c->device, comp);
    if (!P->ok) throw std::runtime_error("implicit diffusion on the MG-PCG path: " + P->why);
    finish_mg_plan(c, *P);
    f.mg = P;
  }
  MGPlan &P = *f.mg;
  P.prec = c->prec == SNAC_PREC_FP64 ? 0 : 1;  // the V-cycle (the network is the pressure's)
  Phase tp(c, SNAC_PH_HELMHOLTZ);
  mg_stage_in(c, P);
  mg_set_shift(P, c->stream, alphai);
  if (c->capturing) split_capture(c, P, mg_solve(P, c->stream, P.fixed_iters));
  else mg_solve(P, c->stream, 0);
  mg_stage_out(c, P);
  for (size_t b = 0; b < c->blk.size(); b++) {
    DevBlock &D = c->blk[b];
    launch_copy_unknowns(c->stream, D.g, f.rhs[b].ext, D.pp, imp_field(c, D, comp));
  }
}

// MG path: block pp interiors -> (transform along the FFT axis) -> plan
// layout scaled by -1 (the solver works with -A) -> MG-PCG -> back
void mg_stage_in(snac_ctx *c, MGPlan &P) {
  const long pst[3] = {1, (long)P.G[0], (long)P.G[0] * P.G[1]};
  const long na = P.G[P.ax_a], ns = P.G[P.ax_s];
  if (P.fa >= 0) {
    int a = P.fa, b1, b2;
    others(a, b1, b2);
    for (size_t b = 0; b < c->blk.size(); b++) {
      DevBlock &D = c->blk[b];
      const Block &B = c->cs.blocks[b];
      const double *in = D.pp + 1 + D.g.sx + D.g.sx * D.g.sy;
      long off = 0;
      for (int q = 0; q < 3; q++) off += (long)(B.lo_g[q] - P.gmin[q]) * pst[q];
      launch_r2r(c->stream, P.fwd, in, gstride(D, a), gstride(D, b1), gstride(D, b2), P.ph + off, pst[a], pst[b1],
                 pst[b2], bext(D, b1), bext(D, b2), 1.0);
    }
    mg_copy3(c->stream, P.ph, pst[P.ax_m], pst[P.ax_a], pst[P.ax_s], P.r, na * ns, ns, 1, P.ng, (int)na,
             (int)ns, -1.0, pst[P.ax_a] == 1 ? 1 : 0);
  } else {
    for (size_t b = 0; b < c->blk.size(); b++) {
      DevBlock &D = c->blk[b];
      const Block &B = c->cs.blocks[b];
      const double *in = D.pp + 1 + D.g.sx + D.g.sx * D.g.sy;
      long off = ((long)(B.lo_g[P.ax_m] - P.gmin[P.ax_m]) * na + (B.lo_g[P.ax_a] - P.gmin[P.ax_a])) * ns +
                 (B.lo_g[P.ax_s] - P.gmin[P.ax_s]);
      mg_copy3(c->stream, in, gstride(D, P.ax_m), gstride(D, P.ax_a), gstride(D, P.ax_s), P.r + off, na * ns, ns, 1,
               B.n[P.ax_m], B.n[P.ax_a], B.n[P.ax_s], -1.0, P.ax_a == 0 ? 1 : 0);
    }
  }
}

void mg_stage_out(snac_ctx *c, MGPlan &P) {
  const long pst[3] = {1, (long)P.G[0], (long)P.G[0] * P.G[1]};
  const long na = P.G[P.ax_a], ns = P.G[P.ax_s];
  if (P.fa >= 0) {
    mg_copy3(c->stream, P.x, na * ns, ns, 1, P.ph, pst[P.ax_m], pst[P.ax_a], pst[P.ax_s], P.ng, (int)na,
             (int)ns, 1.0, pst[P.ax_a] == 1 ? 2 : 0);
    int a = P.fa, b1, b2;
    others(a, b1, b2);
    for (size_t b = 0; b < c->blk.size(); b++) {
      DevBlock &D = c->blk[b];
      const Block &B = c->cs.blocks[b];
      double *out = D.pp + 1 + D.g.sx + D.g.sx * D.g.sy;
      long off = 0;
      for (int q = 0; q < 3; q++) off += (long)(B.lo_g[q] - P.gmin[q]) * pst[q];
      launch_r2r(c->stream, P.bwd, P.ph + off, pst[a], pst[b1], pst[b2], out, gstride(D, a), gstride(D, b1),
                 gstride(D, b2), bext(D, b1), bext(D, b2), P.norm);
    }
  } else {
    for (size_t b = 0; b < c->blk.size(); b++) {
      DevBlock &D = c->blk[b];
      const Block &B = c->cs.blocks[b];
      double *out = D.pp + 1 + D.g.sx + D.g.sx * D.g.sy;
      long off = ((long)(B.lo_g[P.ax_m] - P.gmin[P.ax_m]) * na + (B.lo_g[P.ax_a] - P.gmin[P.ax_a])) * ns +
                 (B.lo_g[P.ax_s] - P.gmin[P.ax_s]);
      mg_copy3(c->stream, P.x + off, na * ns, ns, 1, out, gstride(D, P.ax_m), gstride(D, P.ax_a), gstride(D, P.ax_s),
               B.n[P.ax_m], B.n[P.ax_a], B.n[P.ax_s], 1.0, P.ax_a == 0 ? 2 : 0);
    }
  }
}

// ends the graph segment being captured at the end of a solve's captured
// CG iterations, keeps the solve's state for the host's resumption, and
// starts capturing the next segment
void split_capture(snac_ctx *c, MGPlan &P, const MGRun &R) {
  hipGraph_t g;
  SNAC_HIP(hipStreamEndCapture(c->stream, &g));
  hipGraphExec_t e = nullptr;
  const hipError_t err = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  SNAC_HIP(err);
  c->gseg[c->cap_key].push_back(e);
  c->gres[c->cap_key].push_back({&P, R});
  SNAC_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
}

void solve_pressure(snac_ctx *c) {
  check_solver(c);
  if (use_mg(c)) {
    { Phase tf(c, SNAC_PH_R2R_FWD); mg_stage_in(c, c->mg); }
    {
      Phase tt(c, SNAC_PH_TRIDIAG);
      if (c->capturing) split_capture(c, c->mg, mg_solve(c->mg, c->stream, c->mg.fixed_iters));
      else mg_solve(c->mg, c->stream, 0);
    }
    { Phase tb(c, SNAC_PH_R2R_BWD); mg_stage_out(c, c->mg); }
    return;
  }
  tensor_solve(c, c->tp.P, -1, 0.0);
}

// Autotuner: times every candidate tile shape of each transform call site
// and both tridiagonal variants on this case's own arrays, keeps the
// fastest.  Runs once at context creation.
template <class F>
double time_stage(snac_ctx *c, F &&f) {
  hipEvent_t e0, e1;
  SNAC_HIP(hipEventCreate(&e0));
  SNAC_HIP(hipEventCreate(&e1));
  f();  // warm-up
  SNAC_HIP(hipEventRecord(e0, c->stream));
  const int reps = 5;
  for (int r = 0; r < reps; r++) f();
  SNAC_HIP(hipEventRecord(e1, c->stream));
  SNAC_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  SNAC_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return ms / reps;
}

void autotune(snac_ctx *c) {
  TensorPlan &t = c->tp;
  if (!t.ok || c->device < 0) return;
  t.tune_ms.clear();
  auto tune = [&](R2RPlan &pl, auto &&run) {
    if (!pl.fft) return;
    int best = 0;
    double bt = 1e300;
    const int nt = r2r_tile_choices(pl);
    for (int ls = 0; ls < nt; ls++) {
      pl.lshift = ls;
      double ms = time_stage(c, run);
      if (ms < bt) { bt = ms; best = ls; }
    }
    pl.lshift = best;
    t.tune_ms.push_back(bt);
  };
  tune(t.P.fwd[0], [&] { stage_first(c, t.P); });
  for (size_t q = 1; q < t.xaxes.size(); q++) {
    tune(t.P.fwd[q], [&] { stage_inplace(c, t.P, q, true); });
    tune(t.P.bwd[q], [&] { stage_inplace(c, t.P, q, false); });
  }
  tune(t.P.bwd[0], [&] { stage_last(c, t.P, -1); });
  int best = 1;
  double bt = 1e300;
  const int vmax = tridiag_chunks(t.G[t.s], true) > 0 ? 3 : (tridiag_chunks(t.G[t.s]) > 0 ? 2 : 1);
  for (int v = (t.G[t.s] <= 128 ? 0 : 1); v <= vmax; v++) {
    t.tvariant = v;
    double ms = time_stage(c, [&] { stage_tridiag(c, t.P, 0.0); });
    if (ms < bt) { bt = ms; best = v; }
  }
  t.tvariant = best;
  t.tune_ms.push_back(bt);
  // the tuning runs left transformed zeros behind; restore a clean state
  SNAC_HIP(hipMemsetAsync(t.ph, 0, (size_t)t.G[0] * t.G[1] * t.G[2] * sizeof(double), c->stream));
  for (auto &D : c->blk) SNAC_HIP(hipMemsetAsync(D.pp, 0, D.ntot * sizeof(double), c->stream));
  SNAC_HIP(hipStreamSynchronize(c->stream));
}

// Boussinesq buoyancy of scalar 1 when switched on (the reference's
// _BOUSSINESQ_BUOYANCY build, snac_set_buoyancy); skipped when gacc*beta
// vanishes, where the term is exactly zero
void rk_mom(snac_ctx *c, int irk, double dt) {
  const double f1 = RKC[irk - 1][0] * dt, f2 = RKC[irk - 1][1] * dt, f12 = f1 + f2;
  const Case &cs = c->cs;
  const double gb[3] = {cs.gacc[0] * cs.beta, cs.gacc[1] * cs.beta, cs.gacc[2] * cs.beta};
  const bool buoy = c->buoyancy && cs.nscal > 0 && (gb[0] != 0.0 || gb[1] != 0.0 || gb[2] != 0.0);
  int nx = 1 - c->cur;
  for (auto &D : c->blk) {
    if (c->st32)
      launch_rk_mom_f32(c->stream, D.g, D.m, D.vel[c->cur][0], D.vel[c->cur][1], D.vel[c->cur][2], D.p,
                        D.vel[nx][0], D.vel[nx][1], D.vel[nx][2], D.dto32, D.dto32 + D.g.nx, D.dto32 + 2 * D.g.nx,
                        f1, f2, f12, cs.visc, cs.bforce[0], cs.bforce[1], cs.bforce[2],
                        buoy ? D.sv[c->scur][0] : nullptr, gb, c->rk32_variant, c->impdiff);
    else
      launch_rk_mom(c->stream, D.g, D.m, D.vel[c->cur][0], D.vel[c->cur][1], D.vel[c->cur][2], D.p, D.vel[nx][0],
                    D.vel[nx][1], D.vel[nx][2], D.dto[0], D.dto[1], D.dto[2], f1, f2, f12, cs.visc, cs.bforce[0],
                    cs.bforce[1], cs.bforce[2], buoy ? D.sv[c->scur][0] : nullptr, gb, c->rk_variant, c->impdiff);
  }
  c->cur = nx;
}

// rk_mom kernel variant (one cell per thread, a k-march with one of 6
// chunk/tile choices, or an LDS-staged k-march with one of 4, with or
// without non-temporal hints), timed on this case's own arrays at context
// creation (the arrays are still zero then; the buffers it writes are
// cleared afterwards)
void autotune_rk(snac_ctx *c) {
  if (c->device < 0) return;
  double best = 1e300;
  int bv = 0;
  for (int v = 0; v <= 16; v++) {
    c->rk_variant = v;
    const int cur0 = c->cur;
    double ms = time_stage(c, [&] {
      rk_mom(c, 1, 1e-3);
      c->cur = cur0;
    });
    if (ms < best) { best = ms; bv = v; }
  }
  c->rk_variant = bv;
  c->rk_tune_ms = best;
  // the FP32 form (mixed precision) among the LDS variants
  best = 1e300;
  bv = 16;
  c->st32 = true;
  for (int v = 7; v <= 16; v++) {
    c->rk32_variant = v;
    const int cur0 = c->cur;
    double ms = time_stage(c, [&] {
      rk_mom(c, 1, 1e-3);
      c->cur = cur0;
    });
    if (ms < best) { best = ms; bv = v; }
  }
  c->st32 = false;
  c->rk32_variant = bv;
  c->rk32_tune_ms = best;
  for (auto &D : c->blk) {
    SNAC_HIP(hipMemsetAsync(D.dto32, 0, 3 * D.nin * sizeof(float), c->stream));
    for (int par = 0; par < 2; par++)
      for (int q = 0; q < 3; q++) SNAC_HIP(hipMemsetAsync(D.vel[par][q] - 15, 0, (D.ntot + 16) * sizeof(double), c->stream));
    SNAC_HIP(hipMemsetAsync(D.dto[0], 0, 3 * D.nin * sizeof(double), c->stream));
  }
  SNAC_HIP(hipStreamSynchronize(c->stream));
}

// rk_scal + bulk_forcing_s of every scalar (main.f90:908-913): the update
// goes to the other buffer of each scalar's pair; the forced ones then get
// scalf minus their volume mean over all blocks, computed on the device
void rk_scal(snac_ctx *c, int irk, double dt) {
  const Case &cs = c->cs;
  if (cs.nscal == 0) return;
  const double f1 = RKC[irk - 1][0] * dt, f2 = RKC[irk - 1][1] * dt, f12 = f1 + f2;
  const int nx = 1 - c->scur;
  for (int is = 0; is < cs.nscal; is++) {
    // implicit diffusion: an unforced scalar takes its half diffusion off in
    // the kernel, a forced one after its bulk mean (rk.f90:183-202)
    const int imp = c->impdiff ? (cs.is_sforced[is] ? 2 : 1) : 0;
    for (auto &D : c->blk)
      launch_rk_scal(c->stream, D.g, D.m, D.vel[c->cur][0], D.vel[c->cur][1], D.vel[c->cur][2], D.sv[c->scur][is],
                     D.sv[nx][is], D.sdo[is], cs.alpha[is], f1, f2, f12, cs.ssource[is], imp, D.sdd);
    if (!cs.is_sforced[is]) continue;
    for (auto &D : c->blk) launch_volsum_part(c->stream, D.g, D.m, D.sv[nx][is], c->part_sum + D.part_off);
    launch_reduce_final(c->stream, c->part_sum, nullptr, c->nparts, c->sred);
    for (size_t b = 0; b < c->blk.size(); b++)
      launch_bulk_force(c->stream, c->blk[b].g, c->blk[b].sv[nx][is], c->sred, cs.scalf[is], 1.0 / cs.vol_all,
                        b == 0 ? c->sforce + is : nullptr, imp == 2 ? c->blk[b].sdd : nullptr, 0.5 * f12);
  }
  c->scur = nx;
}

void fillps(snac_ctx *c, double dtrk) {
  for (auto &D : c->blk)
    launch_fillps(c->stream, D.g, D.m, c->cs.fft_axis, D.vel[c->cur][0], D.vel[c->cur][1], D.vel[c->cur][2], D.pp,
                  dtrk, D.rhsp);
}

void correc(snac_ctx *c, double dtrk, bool fuse_pressure) {
  int nx = 1 - c->cur;
  for (auto &D : c->blk)
    launch_correc(c->stream, D.g, D.m, D.pp, D.vel[c->cur][0], D.vel[c->cur][1], D.vel[c->cur][2], D.vel[nx][0],
                  D.vel[nx][1], D.vel[nx][2], dtrk, fuse_pressure ? D.p : nullptr);
  c->cur = nx;
}

void substage(snac_ctx *c, int irk, double dt) {
  double dtrk = (RKC[irk - 1][0] + RKC[irk - 1][1]) * dt;
  if (c->cs.nscal > 0) {
    { Phase t(c, SNAC_PH_RK_SCAL); rk_scal(c, irk, dt); }
    if (c->impdiff)  // main.f90:914-933
      for (int is = 0; is < c->cs.nscal; is++) solve_impdiff(c, 3 + is, -c->cs.alpha[is] * dtrk / 2.0);
    { Phase t(c, SNAC_PH_BOUND); run_program(c, c->prog_scal[c->scur]); }  // boundp of the scalars
  }
  { Phase t(c, SNAC_PH_RK_MOM); rk_mom(c, irk, dt); }
  const double alpha = -c->cs.visc * dtrk / 2.0;  // main.f90:935
  if (c->impdiff)  // main.f90:947-997
    for (int comp = 0; comp < 3; comp++) solve_impdiff(c, comp, alpha);
  { Phase t(c, SNAC_PH_BOUND); run_program(c, c->prog_uvw[c->cur][0]); }  // bounduvw + inflow
  { Phase t(c, SNAC_PH_FILLPS); fillps(c, dtrk); }
  solve_pressure(c);
  { Phase t(c, SNAC_PH_BOUND); run_program(c, c->prog_pp); }               // boundp(pp)
  {
    Phase t(c, SNAC_PH_CORREC);
    correc(c, dtrk, !c->impdiff);  // correc + updt_pressure (p += pp)
    if (c->impdiff)                // updt_pressure with the Laplacian term
      for (auto &D : c->blk) launch_updt_pressure_imp(c->stream, D.g, D.m, D.pp, D.p, alpha);
  }
  { Phase t(c, SNAC_PH_BOUND); run_program(c, c->prog_end[c->cur]); }  // bounduvw/inflow(correc) + boundp(p)
  if (c->pending.size() > 4096) collect_timing(c);  // bounds the event pool in long timed runs
}

// destroys the captured step graphs of one starting parity (all: key < 0);
// a launched graph may still be running on the context's stream, so the
// stream is drained first
void drop_graph(snac_ctx *c, int key = -1) {
  bool any = false;
  for (int q = 0; q < 2; q++) any = any || ((key < 0 || key == q) && !c->gseg[q].empty());
  if (!any) return;
  if (c->stream) SNAC_HIP(hipStreamSynchronize(c->stream));
  for (int q = 0; q < 2; q++) {
    if (key >= 0 && key != q) continue;
    for (auto e : c->gseg[q]) (void)hipGraphExecDestroy(e);
    c->gseg[q].clear();
    c->gres[q].clear();
  }
}

// captures one RK3 step into the segments of key (not run)
void capture_step(snac_ctx *c, int key, double dt) {
  const int cur0 = c->cur;
  c->capturing = true;
  c->cap_key = key;
  SNAC_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  try {
    for (int irk = 1; irk <= 3; irk++) substage(c, irk, dt);
    hipGraph_t g;
    SNAC_HIP(hipStreamEndCapture(c->stream, &g));
    c->capturing = false;
    hipGraphExec_t e = nullptr;
    const hipError_t err = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    SNAC_HIP(err);
    c->gseg[key].push_back(e);
  } catch (...) {
    if (c->capturing) {
      hipGraph_t g = nullptr;
      if (hipStreamEndCapture(c->stream, &g) == hipSuccess && g) (void)hipGraphDestroy(g);
      c->capturing = false;
    }
    for (auto e : c->gseg[key]) (void)hipGraphExecDestroy(e);
    c->gseg[key].clear();
    c->gres[key].clear();
    c->cur = cur0;
    c->scur = key;
    throw;
  }
  if (c->cur != cur0) throw std::runtime_error("internal: a step must return to the same buffer parity");
  c->scur = key;  // captured, not run yet
  c->graph_dt[key] = dt;
  c->graph_cur[key] = c->cur;
}

// nsteps RK3 steps.  Eagerly when timing (events bracket every phase) or
// when graphs are off; otherwise one step is captured into hipGraphs (the
// kernels of a step address the same buffers every step: the velocity
// parity flips six times per step) and replayed, re-captured when dt or the
// starting parity changes.  On the MG-PCG path each solve's first
// fixed_iters CG iterations are part of the graph (the pressure's, and with
// implicit diffusion each field's Helmholtz solve, by its own plan's count);
// the host then resumes the solve until every mode meets hypre_tol (or
// hypre_maxiter), so a graph step ends every solve exactly where an eager
// one does.
void run_steps(snac_ctx *c, double dt, int nsteps) {
  check_solver(c);
  if (nsteps <= 0) return;
  // the learned preconditioner and the pair recording decide on the host
  // between CG iterations: eager launches
  const bool host_cg = use_mg(c) && (c->prec == SNAC_PREC_LEARNED || c->mg.collecting);
  if (!c->use_graph || c->timing || host_cg) {
    for (int s = 0; s < nsteps; s++)
      for (int irk = 1; irk <= 3; irk++) substage(c, irk, dt);
    return;
  }
  const bool mgp = use_mg(c);
  MGPlan &P = c->mg;
  // the Helmholtz solves of implicit diffusion on the MG-PCG path (their
  // plans exist after the first step that solves them)
  bool imp_new = false;
  std::vector<MGPlan *> hp;
  if (mgp && c->impdiff)
    for (auto &f : c->imp) {
      if (!f.mg_ok && c->tp.ok) continue;  // solved on the tensor-product plan
      if (!f.mg || f.mg->fixed_iters == 0) imp_new = true;
      if (f.mg) hp.push_back(f.mg.get());
    }
  if (mgp && (P.fixed_iters == 0 || imp_new)) {
    // one step with host-checked convergence measures the CG count the
    // captured part of each solve runs
    for (int irk = 1; irk <= 3; irk++) substage(c, irk, dt);
    nsteps--;
    P.fixed_iters = c->force_fixed > 0 ? std::min(c->force_fixed, P.maxiter)
                                       : std::min(P.maxiter, std::max(1, P.max_seen + 1));
    hp.clear();
    for (auto &f : c->imp)
      if (f.mg) {
        f.mg->fixed_iters = std::min(f.mg->maxiter, std::max(1, f.mg->max_seen + 1));
        hp.push_back(f.mg.get());
      }
    drop_graph(c);
  }
  // a step flips the velocity parity six times (back to the start) and the
  // scalar parity three times: one graph per starting scalar parity
  for (int s = 0; s < nsteps; s++) {
    const int key = c->scur;
    if (c->gseg[key].empty() || c->graph_dt[key] != dt || c->graph_cur[key] != c->cur) {
      drop_graph(c, key);
      capture_step(c, key, dt);
    }
    const size_t nseg = c->gseg[key].size();
    for (size_t q = 0; q < nseg; q++) {
      SNAC_HIP(hipGraphLaunch(c->gseg[key][q], c->stream));
      if (q + 1 < nseg) {
        MGRun R = c->gres[key][q].second;
        mg_continue(*c->gres[key][q].first, c->stream, R);
      }
    }
    if (c->cs.nscal > 0) c->scur = 1 - key;
  }
  // solves that needed more than the captured count raise it for the next
  // capture (the count only grows, up to maxiter)
  bool grow = false;
  if (mgp && c->force_fixed == 0 && P.max_seen + 1 > P.fixed_iters && P.fixed_iters < P.maxiter) {
    P.fixed_iters = std::min(P.maxiter, P.max_seen + 1);
    grow = true;
  }
  for (MGPlan *H : hp)
    if (H->max_seen + 1 > H->fixed_iters && H->fixed_iters < H->maxiter) {
      H->fixed_iters = std::min(H->maxiter, H->max_seen + 1);
      grow = true;
    }
  if (grow) drop_graph(c);
}

double *field_ptr(snac_ctx *c, int b, int fid, long &count) {
  DevBlock &D = c->blk.at(b);
  if (fid >= 0 && fid < 6) count = D.ntot;
  else count = D.nin;
  switch (fid) {
  case SNAC_U: return D.vel[c->cur][0];
  case SNAC_V: return D.vel[c->cur][1];
  case SNAC_W: return D.vel[c->cur][2];
  case SNAC_P: return D.p;
  case SNAC_PP: return D.pp;
  case SNAC_PO: return D.po;
  case SNAC_DUDTRKO: return D.dto[0];
  case SNAC_DVDTRKO: return D.dto[1];
  case SNAC_DWDTRKO: return D.dto[2];
  default: break;
  }
  if (fid >= SNAC_SCAL0 && fid < SNAC_SCAL0 + 2 * c->cs.nscal) {
    const int is = (fid - SNAC_SCAL0) / 2;
    if ((fid - SNAC_SCAL0) % 2 == 0) { count = D.ntot; return D.sv[c->scur][is]; }
    count = D.nin;
    return D.sdo[is];
  }
  throw std::runtime_error("invalid field id " + std::to_string(fid));
}

void reduce(snac_ctx *c, bool div, double out[2]) {
  for (auto &D : c->blk) {
    if (div)
      launch_chkdiv_part(c->stream, D.g, D.m, D.vel[c->cur][0], D.vel[c->cur][1], D.vel[c->cur][2], c->cs.vol_all,
                         c->part_sum + D.part_off, c->part_max + D.part_off);
    else
      launch_chkdt_part(c->stream, D.g, D.m, D.vel[c->cur][0], D.vel[c->cur][1], D.vel[c->cur][2],
                        c->part_max + D.part_off);
  }
  launch_reduce_final(c->stream, div ? c->part_sum : nullptr, c->part_max, c->nparts, c->red);
  SNAC_HIP(hipMemcpyAsync(out, c->red, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  SNAC_HIP(hipStreamSynchronize(c->stream));
}

// The autotuner's choices per problem shape, kept for the life of the
// process: a second context of the same shapes runs the same kernels
// without tuning again, so the two agree bit for bit (tiles that factor a
// transform differently round differently).
struct TuneChoice {
  std::vector<int> fwd, bwd;
  int tvariant = 1, rk_variant = 0, rk32_variant = 16;
  std::vector<double> tune_ms;
  double rk_ms = 0.0, rk32_ms = 0.0;
};
std::mutex g_tune_mu;
std::map<std::string, TuneChoice> g_tune;

std::string tune_key(const snac_ctx *c) {
  std::ostringstream k;
  k << "dev " << c->device << " scal " << c->cs.nscal << " blocks";
  for (const auto &D : c->blk) k << ' ' << D.g.nx << 'x' << D.g.ny << 'x' << D.g.nz;
  const TensorPlan &t = c->tp;
  if (t.ok) {
    k << " tensor " << t.G[0] << 'x' << t.G[1] << 'x' << t.G[2] << " line " << t.s << " axes";
    for (int a : t.xaxes) k << ' ' << a;
    for (const auto &p : t.P.fwd) k << " f" << p.kind << ':' << p.n;
    for (const auto &p : t.P.bwd) k << " b" << p.kind << ':' << p.n;
  }
  return k.str();
}

void tune_or_recall(snac_ctx *c) {
  if (c->device < 0) return;
  const std::string key = tune_key(c);
  std::lock_guard<std::mutex> lk(g_tune_mu);
  TensorPlan &t = c->tp;
  auto it = g_tune.find(key);
  if (it != g_tune.end()) {
    const TuneChoice &tc = it->second;
    if (t.ok) {
      for (size_t q = 0; q < t.P.fwd.size(); q++) t.P.fwd[q].lshift = tc.fwd[q];
      for (size_t q = 0; q < t.P.bwd.size(); q++) t.P.bwd[q].lshift = tc.bwd[q];
      t.tvariant = tc.tvariant;
      t.tune_ms = tc.tune_ms;
    }
    c->rk_variant = tc.rk_variant;
    c->rk_tune_ms = tc.rk_ms;
    c->rk32_variant = tc.rk32_variant;
    c->rk32_tune_ms = tc.rk32_ms;
    return;
  }
  autotune(c);
  autotune_rk(c);
  TuneChoice tc;
  if (t.ok) {
    for (const auto &p : t.P.fwd) tc.fwd.push_back(p.lshift);
    for (const auto &p : t.P.bwd) tc.bwd.push_back(p.lshift);
    tc.tvariant = t.tvariant;
    tc.tune_ms = t.tune_ms;
  }
  tc.rk_variant = c->rk_variant;
  tc.rk_ms = c->rk_tune_ms;
  tc.rk32_variant = c->rk32_variant;
  tc.rk32_ms = c->rk32_tune_ms;
  g_tune.emplace(key, tc);
}

void create(snac_ctx *c) {
  if (c->device >= 0) {
    SNAC_HIP(hipSetDevice(c->device));
    SNAC_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  }
  const Case &cs = c->cs;
  c->blk.resize(cs.blocks.size());
  c->dlmin = 1e300;
  for (size_t b = 0; b < cs.blocks.size(); b++) {
    const Block &B = cs.blocks[b];
    DevBlock &D = c->blk[b];
    D.g.nx = B.n[0]; D.g.ny = B.n[1]; D.g.nz = B.n[2];
    D.g.sx = ((B.n[0] + 2 + 15) / 16) * 16;  // padded leading dimension
    D.g.sy = B.n[1] + 2;
    D.ntot = D.g.sx * D.g.sy * (long)(B.n[2] + 2);
    D.nin = (long)B.n[0] * B.n[1] * B.n[2];
    for (int a = 0; a < 3; a++) {
      D.met[2 * a] = dupload(c, B.dc[a]);
      D.met[2 * a + 1] = dupload(c, B.df[a]);
      std::vector<double> ic(B.dc[a].size()), iff(B.df[a].size());
      for (size_t q = 0; q < ic.size(); q++) { ic[q] = 1.0 / B.dc[a][q]; iff[q] = 1.0 / B.df[a][q]; }
      D.met[6 + 2 * a] = dupload(c, ic);
      D.met[6 + 2 * a + 1] = dupload(c, iff);
      for (double v : B.df[a]) c->dlmin = std::min(c->dlmin, v);
    }
    D.m = Metrics{D.met[0], D.met[1], D.met[2], D.met[3], D.met[4], D.met[5],
                  D.met[6], D.met[7], D.met[8], D.met[9], D.met[10], D.met[11]};
    for (int par = 0; par < 2; par++)
      for (int q = 0; q < 3; q++) D.vel[par][q] = dalloc_field(c, D.ntot);
    D.p = dalloc_field(c, D.ntot);
    D.pp = dalloc_field(c, D.ntot);
    D.po = dalloc_field(c, D.ntot);
    // dudtrko, dvdtrko, dwdtrko row by row in one array ([k][j][component][i])
    D.dto[0] = dalloc<double>(c, 3 * D.nin);
    for (int q = 1; q < 3; q++) D.dto[q] = D.dto[0] + q * (long)D.g.nx;
    D.dto32 = dalloc<float>(c, 3 * D.nin);
    for (int is = 0; is < cs.nscal; is++) {
      for (int par = 0; par < 2; par++) D.sv[par].push_back(dalloc_field(c, D.ntot));
      D.sdo.push_back(dalloc<double>(c, D.nin));
    }
    for (int a = 0; a < 3; a++)
      for (int s = 0; s < 2; s++) {
        D.rhsp[2 * a + s] = B.rhsp[a][s].empty() ? nullptr : dupload(c, B.rhsp[a][s]);
        for (int q = 0; q < 3; q++) D.vin[a][q][s] = B.vin[a][q][s].empty() ? nullptr : dupload(c, B.vin[a][q][s]);
      }
    D.part_off = c->nparts;
    D.nparts = reduce_nparts(D.g);
    c->nparts += D.nparts;
  }
  c->part_sum = dalloc<double>(c, c->nparts);
  c->part_max = dalloc<double>(c, c->nparts);
  c->red = dalloc<double>(c, 2);
  auto arrs = block_arrays(c);
  for (int par = 0; par < 2; par++)
    for (int cr = 0; cr < 2; cr++) {
      c->prog_uvw[par][cr] = upload_program(c, build_velocity_program(cs, arrs, par, cr, true, true));
      c->prog_bound_only[par][cr] = upload_program(c, build_velocity_program(cs, arrs, par, cr, true, false));
      c->prog_inflow[par][cr] = upload_program(c, build_velocity_program(cs, arrs, par, cr, false, true));
    }
  for (int par = 0; par < 2; par++) c->prog_end[par] = upload_program(c, build_correc_fill_program(cs, arrs, par));
  c->prog_p = upload_program(c, build_pressure_program(cs, arrs, 0));
  c->prog_pp = upload_program(c, build_pressure_program(cs, arrs, 1));
  if (cs.nscal > 0) {
    for (int par = 0; par < 2; par++) {
      c->prog_scal[par] = upload_program(c, build_scalar_program(cs, arrs, par, -1));
      for (int is = 0; is < cs.nscal; is++)
        c->prog_scal_one[par].push_back(upload_program(c, build_scalar_program(cs, arrs, par, is)));
    }
    c->sred = dalloc<double>(c, 2);
    c->sforce = dalloc<double>(c, cs.nscal);
  }
  plan_tensor(c);
  tune_or_recall(c);
  if (!c->tp.ok) plan_mg(c);
  if (c->device >= 0) SNAC_HIP(hipDeviceSynchronize());
}

void destroy(snac_ctx *c) {
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  mg_free(c->mg);
  drop_graph(c);
  for (auto &r : c->pending) { (void)hipEventDestroy(r.a); (void)hipEventDestroy(r.b); }
  for (auto &e : c->pool) (void)hipEventDestroy(e);
  for (void *p : c->owned) (void)hipFree(p);
  if (c->stream) (void)hipStreamDestroy(c->stream);
}

template <class F>
int guard(F &&f) {
  try {
    g_err.clear();
    f();
    return 0;
  } catch (const std::exception &e) {
    g_err = e.what();
    return -1;
  } catch (...) {
    g_err = "unknown error";
    return -1;
  }
}

snac_ctx *checked(snac_ctx *c) {
  if (!c) throw std::runtime_error("null context");
  return c;
}

// compute entry points need a device context
snac_ctx *on_device(snac_ctx *c) {
  checked(c);
  if (c->device < 0) throw std::runtime_error("host-only context (device < 0): set-up and queries only");
  return c;
}

}  // namespace
}  // namespace snac

using namespace snac;

namespace snac {
// the message snac_last_error returns on this host thread
void set_last_error(const std::string &msg) { g_err = msg; }
}  // namespace snac

extern "C" {

const char *snac_last_error(void) { return g_err.c_str(); }

int snac_ctx_create_from_nml(snac_ctx **out, const char *dns, const char *blocks, int fft_axis, int device) {
  return guard([&] {
    if (!out || !dns || !blocks) throw std::runtime_error("null argument");
    *out = nullptr;
    std::unique_ptr<snac_ctx> c(new snac_ctx());
    c->cs = build_case(dns, blocks, fft_axis);
    c->device = device;
    try {
      create(c.get());
    } catch (...) {
      destroy(c.get());
      throw;
    }
    *out = c.release();
  });
}

int snac_ctx_destroy(snac_ctx *c) {
  return guard([&] {
    if (!c) return;
    destroy(c);
    delete c;
  });
}

int snac_ctx_num_blocks(const snac_ctx *c) { return c ? (int)c->cs.blocks.size() : -1; }

int snac_ctx_block_info(const snac_ctx *c, int b, int *n3, int *lo3, int *nb6, int *per3) {
  return guard([&] {
    const Block &B = checked(const_cast<snac_ctx *>(c))->cs.blocks.at(b);
    for (int a = 0; a < 3; a++) {
      if (n3) n3[a] = B.n[a];
      if (lo3) lo3[a] = B.lo_g[a];
      if (nb6) { nb6[2 * a] = B.nb[a][0]; nb6[2 * a + 1] = B.nb[a][1]; }
      if (per3) per3[a] = c->cs.periods[a];
    }
  });
}

int snac_ctx_get_metric(const snac_ctx *c, int b, int which, double *out) {
  return guard([&] {
    const Block &B = checked(const_cast<snac_ctx *>(c))->cs.blocks.at(b);
    if (which < 0 || which > 11) throw std::runtime_error("invalid metric id");
    int a = (which % 6) / 2;
    const std::vector<double> &v = which < 6 ? ((which % 2) ? B.df[a] : B.dc[a]) : ((which % 2) ? B.rf[a] : B.rc[a]);
    std::memcpy(out, v.data(), v.size() * sizeof(double));
  });
}

int snac_ctx_get_params(const snac_ctx *c, double *dp, int *ip) {
  return guard([&] {
    const Case &s = checked(const_cast<snac_ctx *>(c))->cs;
    if (dp) {
      double v[11] = {s.visc, s.cfl, s.dtmax, s.dt_f, s.bforce[0], s.bforce[1], s.bforce[2], s.vol_all, s.time_max,
                      s.tw_max, s.hypre_tol};
      std::memcpy(dp, v, sizeof(v));
    }
    if (ip) {
      int v[8] = {s.nstep, s.icheck, s.nscal, s.hypre_maxiter, (int)s.restart, (int)s.stop_type[0],
                  (int)s.stop_type[1], (int)s.stop_type[2]};
      std::memcpy(ip, v, sizeof(v));
    }
  });
}

int snac_solver_stats(snac_ctx *c, int *iters, double *max_relres) {
  return guard([&] {
    on_device(c);
    int it = 0;
    double rel = 0.0;
    if (use_mg(c) && c->mg.ok) mg_stats(c->mg, c->stream, it, rel);
    if (iters) *iters = it;
    if (max_relres) *max_relres = rel;
  });
}

int snac_solver_counters(snac_ctx *c, int *solves, int *continued, int *unconverged) {
  return guard([&] {
    checked(c);
    MGPlan &P = c->mg;
    if (solves) *solves = P.n_solves;
    if (continued) *continued = P.n_continued;
    if (unconverged) *unconverged = P.n_unconverged;
    P.n_solves = P.n_continued = P.n_unconverged = 0;
  });
}

int snac_set_graph_cg_iters(snac_ctx *c, int n) {
  return guard([&] {
    checked(c);
    if (n < 0) throw std::runtime_error("graph CG iteration count must be >= 0");
    drop_graph(c);
    c->force_fixed = n;
    c->mg.fixed_iters = 0;  // measured (or set) again on the next graph step
  });
}

int snac_set_precision(snac_ctx *c, int mode) {
  return guard([&] {
    checked(c);
    if (mode != SNAC_PREC_FP64 && mode != SNAC_PREC_MIXED && mode != SNAC_PREC_LEARNED)
      throw std::runtime_error("unknown precision mode");
    if (mode == SNAC_PREC_LEARNED && !(c->mg_planned && c->mg.net))
      throw std::runtime_error("the learned preconditioner needs a trained network (snac_learned_train)");
    if (c->device >= 0) SNAC_HIP(hipStreamSynchronize(c->stream));
    // the momentum right-hand side runs in FP32 in the mixed mode (the
    // learned mode changes only the preconditioner); the old tendencies move
    // to the other precision's array
    const bool st32 = mode == SNAC_PREC_MIXED;
    if (st32 != c->st32 && c->device >= 0) {
      for (auto &D : c->blk) {
        if (st32) launch_cvt_d2f(c->stream, D.dto[0], D.dto32, 3 * D.nin);
        else launch_cvt_f2d(c->stream, D.dto32, D.dto[0], 3 * D.nin);
      }
      SNAC_HIP(hipStreamSynchronize(c->stream));
    }
    c->st32 = st32;
    c->prec = mode;
    c->mg.prec = mode;
    c->mg.fixed_iters = 0;  // the graph's CG counts are measured again
    c->mg.max_seen = 0;
    for (auto &f : c->imp)
      if (f.mg) {
        f.mg->fixed_iters = 0;
        f.mg->max_seen = 0;
      }
    drop_graph(c);
  });
}

int snac_learned_collect(snac_ctx *c, int npairs) {
  return guard([&] {
    on_device(c);
    check_solver(c);
    MGPlan &P = c->mg;
    if (!use_mg(c) || !c->mg_planned || !P.ok || P.fa < 0)
      throw std::runtime_error("the learned preconditioner needs the MG-PCG path with an FFT axis");
    if (npairs < P.ng) throw std::runtime_error("collect at least one solve's modes (" + std::to_string(P.ng) + ")");
    SNAC_HIP(hipStreamSynchronize(c->stream));
    const size_t pl = (size_t)P.nH * P.nW;
    if (npairs > P.pair_cap) {
      P.pair_x = dalloc<float>(c, (size_t)npairs * pl * 2);
      P.pair_t = dalloc<float>(c, (size_t)npairs * pl);
      P.pair_cap = npairs;
    }
    P.pair_n = 0;
    P.collecting = true;
    drop_graph(c);
  });
}

int snac_learned_train(snac_ctx *c, int epochs, int batch, float lr, unsigned seed, float *loss_per_epoch) {
  return guard([&] {
    on_device(c);
    MGPlan &P = c->mg;
    if (!c->mg_planned || P.pair_n < 1) throw std::runtime_error("no recorded pairs (snac_learned_collect)");
    if (epochs < 1 || batch < 1 || batch > P.pair_n) throw std::runtime_error("learned_train: invalid sizes");
    SNAC_HIP(hipStreamSynchronize(c->stream));
    P.collecting = false;
    if (!P.net) P.net = new Net(P.nH, P.nW, std::max(batch, P.ng), seed);
    if (batch > P.net->maxbatch()) throw std::runtime_error("learned_train: batch above the network's");
    if (!P.train_loss) P.train_loss = dalloc<float>(c, 1);
    float *loss = P.train_loss;
    const size_t pl = (size_t)P.nH * P.nW;
    for (int e = 0; e < epochs; e++) {
      double sum = 0.0;
      int nb = 0;
      for (int s0 = 0; s0 + batch <= P.pair_n; s0 += batch, nb++) {
        P.net->set_input2(c->stream, batch, P.pair_x + (size_t)s0 * pl * 2);
        P.net->forward(c->stream, batch);
        P.net->loss_backward(c->stream, batch, P.pair_t + (size_t)s0 * pl, loss);
        P.net->adam_step(c->stream, lr);
        float l = 0.f;
        SNAC_HIP(hipMemcpyAsync(&l, loss, sizeof(float), hipMemcpyDeviceToHost, c->stream));
        SNAC_HIP(hipStreamSynchronize(c->stream));
        sum += l;
      }
      if (loss_per_epoch) loss_per_epoch[e] = (float)(sum / std::max(nb, 1));
    }
  });
}

int snac_learned_options(snac_ctx *c, int fallback, double stall_ratio, int initial_guess) {
  return guard([&] {
    checked(c);
    if (!(stall_ratio > 0.0 && stall_ratio <= 1.0)) throw std::runtime_error("stall ratio must be in (0, 1]");
    c->mg.fallback = fallback != 0;
    c->mg.stall_ratio = stall_ratio;
    c->mg.guess = initial_guess != 0;
  });
}

int snac_learned_time(snac_ctx *c, int reps, double *ms, double *gflop) {
  return guard([&] {
    on_device(c);
    MGPlan &P = c->mg;
    if (!c->mg_planned || !P.net) throw std::runtime_error("no trained network");
    if (reps < 1) throw std::runtime_error("reps must be positive");
    hipEvent_t e0, e1;
    SNAC_HIP(hipEventCreate(&e0));
    SNAC_HIP(hipEventCreate(&e1));
    P.net->forward(c->stream, P.ng);  // warm-up
    SNAC_HIP(hipEventRecord(e0, c->stream));
    for (int r = 0; r < reps; r++) P.net->forward(c->stream, P.ng);
    SNAC_HIP(hipEventRecord(e1, c->stream));
    SNAC_HIP(hipEventSynchronize(e1));
    float t = 0.f;
    SNAC_HIP(hipEventElapsedTime(&t, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (ms) *ms = t / reps;
    // the convolutions' multiply-adds as the MFMA runs them (padded channels)
    double f = 0.0;
    for (int l = 0; l < NET_LAYERS; l++)
      f += 2.0 * 9.0 * NET_CIN[l] * NET_COUT[l] * ((double)(P.nH >> NET_LEVEL[l]) * (P.nW >> NET_LEVEL[l]));
    if (gflop) *gflop = f * P.ng / 1e9;
  });
}

int snac_learned_stats(snac_ctx *c, int *pairs, int *fallbacks) {
  return guard([&] {
    checked(c);
    if (pairs) *pairs = c->mg.pair_n;
    if (fallbacks) *fallbacks = c->mg.fallbacks;
  });
}

int snac_ctx_set_solver(snac_ctx *c, int solver) {
  return guard([&] {
    checked(c);
    if (solver != SNAC_SOLVER_AUTO && solver != SNAC_SOLVER_TENSOR && solver != SNAC_SOLVER_MG)
      throw std::runtime_error("unknown solver id");
    const int old = c->solver;
    c->solver = solver;
    try {
      check_solver(c);
    } catch (...) {
      c->solver = old;
      throw;
    }
    drop_graph(c);
  });
}

// ghosted fields cross the boundary in the reference layout ((nx+2) doubles
// per row) and live on the device with the padded row pitch
void copy_field(snac_ctx *c, int b, int fid, double *dst_host, const double *src_host) {
  long n;
  double *d = field_ptr(c, b, fid, n);
  DevBlock &D = c->blk.at(b);
  const bool ghosted = (fid >= 0 && fid < 6) || (fid >= SNAC_SCAL0 && (fid - SNAC_SCAL0) % 2 == 0);
  if (fid >= SNAC_DUDTRKO && fid <= SNAC_DWDTRKO) {
    // one component's rows, 3 nx apart on the device, contiguous on the host;
    // with FP32 stencils the tendencies live in dto32 and pass through the
    // FP64 array
    size_t w = (size_t)D.g.nx * sizeof(double), pitch = 3 * w, rows = (size_t)D.g.ny * D.g.nz;
    if (c->st32) launch_cvt_f2d(c->stream, D.dto32, D.dto[0], 3 * D.nin);
    if (src_host) SNAC_HIP(hipMemcpy2DAsync(d, pitch, src_host, w, w, rows, hipMemcpyHostToDevice, c->stream));
    else SNAC_HIP(hipMemcpy2DAsync(dst_host, w, d, pitch, w, rows, hipMemcpyDeviceToHost, c->stream));
    if (c->st32 && src_host) launch_cvt_d2f(c->stream, D.dto[0], D.dto32, 3 * D.nin);
  } else if (ghosted) {
    size_t w = (size_t)(D.g.nx + 2) * sizeof(double), pitch = (size_t)D.g.sx * sizeof(double);
    size_t rows = (size_t)(D.g.ny + 2) * (D.g.nz + 2);
    if (src_host) SNAC_HIP(hipMemcpy2DAsync(d, pitch, src_host, w, w, rows, hipMemcpyHostToDevice, c->stream));
    else SNAC_HIP(hipMemcpy2DAsync(dst_host, w, d, pitch, w, rows, hipMemcpyDeviceToHost, c->stream));
  } else {
    if (src_host) SNAC_HIP(hipMemcpyAsync(d, src_host, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    else SNAC_HIP(hipMemcpyAsync(dst_host, d, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  }
  SNAC_HIP(hipStreamSynchronize(c->stream));
}

int snac_upload_field(snac_ctx *c, int b, int fid, const double *h) {
  return guard([&] { copy_field(on_device(c), b, fid, nullptr, h); });
}

int snac_download_field(snac_ctx *c, int b, int fid, double *h) {
  return guard([&] { copy_field(on_device(c), b, fid, h, nullptr); });
}

}  // extern "C"

// ---- files in the reference's formats (io.hpp) ----
namespace snac {
namespace {

bool is_ghosted(int fid) { return (fid >= 0 && fid < 6) || (fid >= SNAC_SCAL0 && (fid - SNAC_SCAL0) % 2 == 0); }

// the block's interior values of a field, Fortran order
std::vector<double> get_interior(snac_ctx *c, int b, int fid) {
  const DevBlock &D = c->blk.at(b);
  const size_t nx = D.g.nx, ny = D.g.ny, nz = D.g.nz;
  std::vector<double> v(nx * ny * nz);
  if (!is_ghosted(fid)) {
    copy_field(c, b, fid, v.data(), nullptr);
    return v;
  }
  std::vector<double> g((nx + 2) * (ny + 2) * (nz + 2));
  copy_field(c, b, fid, g.data(), nullptr);
  for (size_t k = 1; k <= nz; k++)
    for (size_t j = 1; j <= ny; j++)
      std::memcpy(&v[((k - 1) * ny + (j - 1)) * nx], &g[(k * (ny + 2) + j) * (nx + 2) + 1], nx * sizeof(double));
  return v;
}

// sets a field to the given interior values with zero ghosts (the reference
// zeroes the fields before it loads a checkpoint, main.f90:527-530)
void set_interior(snac_ctx *c, int b, int fid, const std::vector<double> &v) {
  const DevBlock &D = c->blk.at(b);
  const size_t nx = D.g.nx, ny = D.g.ny, nz = D.g.nz;
  if (!is_ghosted(fid)) {
    copy_field(c, b, fid, nullptr, v.data());
    return;
  }
  std::vector<double> g((nx + 2) * (ny + 2) * (nz + 2), 0.0);
  for (size_t k = 1; k <= nz; k++)
    for (size_t j = 1; j <= ny; j++)
      std::memcpy(&g[(k * (ny + 2) + j) * (nx + 2) + 1], &v[((k - 1) * ny + (j - 1)) * nx], nx * sizeof(double));
  copy_field(c, b, fid, nullptr, g.data());
}

void load_one(snac_ctx *c, char io, const std::string &path, int b, int fid, double *time, int *istep) {
  const DevBlock &D = c->blk.at(b);
  const size_t n = (size_t)D.g.nx * D.g.ny * D.g.nz;
  if (io == 'w') {
    const std::vector<double> v = get_interior(c, b, fid);
    write_field_file(path, v.data(), n, time ? *time : 0.0, istep ? *istep : 0);
  } else if (io == 'r') {
    std::vector<double> v(n);
    double t;
    int is;
    read_field_file(path, v.data(), n, t, is);
    set_interior(c, b, fid, v);
    if (time) *time = t;
    if (istep) *istep = is;
  } else {
    throw std::runtime_error("load_one: io must be 'r' or 'w'");
  }
}

}  // namespace
}  // namespace snac

extern "C" {

int snac_load_one(snac_ctx *c, char io, const char *filename, int b, int fid, double *time, int *istep) {
  return guard([&] {
    on_device(c);
    if (!filename) throw std::runtime_error("load_one: no file name");
    load_one(c, io, filename, b, fid, time, istep);
  });
}

int snac_restart_io(snac_ctx *c, char io, const char *datadir, double *time, int *istep) {
  return guard([&] {
    on_device(c);
    if (io != 'r' && io != 'w') throw std::runtime_error("restart_io: io must be 'r' or 'w'");
    const std::string dir = datadir ? datadir : "";
    static const char *names[4] = {"u", "v", "w", "p"};
    for (size_t b = 0; b < c->blk.size(); b++) {
      const std::string base = dir + "fld_b_" + block_num((int)b);
      for (int f = 0; f < 4; f++) load_one(c, io, base + "_" + names[f] + ".bin", (int)b, f, time, istep);
      for (int is = 0; is < c->cs.nscal; is++)
        load_one(c, io, base + "_s_" + block_num(is) + ".bin", (int)b, SNAC_SCAL0 + 2 * is, time, istep);
      if (io == 'r') set_interior(c, (int)b, SNAC_PO, get_interior(c, (int)b, SNAC_P));  // po = p
    }
  });
}

int snac_write_visu_3d(snac_ctx *c, const char *datadir, const char *fname_bin, const char *fname_log,
                       const char *varname, int b, int fid, const int *nmin, const int *nmax, const int *nskip,
                       double time, int istep) {
  return guard([&] {
    on_device(c);
    if (!fname_bin || !fname_log || !varname || !nmin || !nmax || !nskip)
      throw std::runtime_error("write_visu_3d: null argument");
    const std::string dir = datadir ? datadir : "";
    const std::vector<double> v = get_interior(c, b, fid);
    write_raw_field(dir + fname_bin, v.data(), v.size());
    append_visu_log(dir + fname_log, fname_bin, varname, nmin, nmax, nskip, time, istep);
  });
}

int snac_initflow(snac_ctx *c) {
  return guard([&] {
    on_device(c);
    const Case &cs = c->cs;
    const double pi = std::acos(-1.0);
    for (size_t b = 0; b < cs.blocks.size(); b++) {
      const Block &B = cs.blocks[b];
      const int nx = B.n[0], ny = B.n[1], nz = B.n[2];
      const double l[3] = {B.lmax[0] - B.lmin[0], B.lmax[1] - B.lmin[1], B.lmax[2] - B.lmin[2]};
      const size_t ntot = (size_t)(nx + 2) * (ny + 2) * (nz + 2);
      auto at = [&](int i, int j, int k) { return ((size_t)k * (ny + 2) + j) * (nx + 2) + i; };
      const std::vector<double> &xc = B.rc[0], &xf = B.rf[0], &yc = B.rc[1], &yf = B.rf[1], &zc = B.rc[2],
                                &zf = B.rf[2];
      std::vector<double> u(ntot, 0.0), v(ntot, 0.0), w(ntot, 0.0), p(ntot, 0.0);
      // initflow (initflow.f90:10-137), is_wallturb = .false. as main.f90:534 calls it
      const std::string iv = B.inivel;
      std::vector<double> u1d(nz + 2, 0.0);
      double uref = 1.0, ubulk = uref;
      const double lref1 = 1.0;
      bool is_mean = false;
      auto poiseuille = [&](double len, double norm) {
        for (int k = 1; k <= nz; k++) {
          const double z = zc[k] / len;
          u1d[k] = 6.0 * z * (1.0 - z) * norm;
        }
      };
      if (iv == "cou") {
        for (int k = 1; k <= nz; k++) {
          const double z = zc[k] / l[2];
          u1d[k] = 0.5 * (1.0 - 2.0 * z) * 1.0 + 0.5;
          u1d[k] = B.bcvel[0][2][0] * u1d[k] + B.bcvel[0][2][1] * (1.0 - u1d[k]);
        }
      } else if (iv == "poi") {
        poiseuille(l[2], ubulk);
        is_mean = true;
      } else if (iv == "zer") {
      } else if (iv == "uni") {
        for (int k = 1; k <= nz; k++) u1d[k] = uref;
      } else if (iv == "hcp") {
        poiseuille(2.0 * l[2], ubulk);
        is_mean = true;
      } else if (iv == "pdc") {
        const double lref = l[2] / 2.0;
        ubulk = cs.bforce[0] * lref * lref / (3.0 * cs.visc);
        poiseuille(l[2], ubulk);
        is_mean = true;
      } else if (iv == "tgv") {
        for (int k = 1; k <= nz; k++)
          for (int j = 1; j <= ny; j++)
            for (int i = 1; i <= nx; i++) {
              const double zcl = zc[k] / l[2] * 2.0 * pi, ycl = yc[j] / l[1] * 2.0 * pi, yfl = yf[j] / l[1] * 2.0 * pi;
              const double xcl = xc[i] / l[0] * 2.0 * pi, xfl = xf[i] / l[0] * 2.0 * pi;
              u[at(i, j, k)] = std::sin(xfl) * std::cos(ycl) * std::cos(zcl);
              v[at(i, j, k)] = -std::cos(xcl) * std::sin(yfl) * std::cos(zcl);
            }
      } else if (iv == "kov") {
        const double reb = uref * lref1 / cs.visc;
        const double lambda = reb / 2 - std::sqrt(reb * reb / 4 + 4 * pi * pi);
        for (int k = 1; k <= nz; k++)
          for (int j = 1; j <= ny; j++)
            for (int i = 1; i <= nx; i++) {
              u[at(i, j, k)] = 1.0 - std::exp(lambda * xf[i]) * std::cos(2 * pi * zc[k]);
              w[at(i, j, k)] = lambda / (2.0 * pi) * std::exp(lambda * xc[i]) * std::sin(2 * pi * zf[k]);
            }
      } else if (iv == "log" || iv == "hcl") {
        throw std::runtime_error("initial velocity field '" + iv +
                                 "' adds noise from the Fortran compiler's random_number; set the fields with "
                                 "snac_upload_field instead");
      } else {
        throw std::runtime_error("ERROR: invalid name for initial velocity field");
      }
      if (iv != "tgv" && iv != "kov")
        for (int k = 1; k <= nz; k++)
          for (int j = 1; j <= ny; j++)
            for (int i = 1; i <= nx; i++) u[at(i, j, k)] = u1d[k];
      if (is_mean) {  // set_mean (initflow.f90:288-317) with dxc, dyf, dzf over the block
        double meanold = 0.0;
        for (int k = 1; k <= nz; k++)
          for (int j = 1; j <= ny; j++)
            for (int i = 1; i <= nx; i++)
              meanold = meanold + u[at(i, j, k)] * B.dc[0][i] * B.df[1][j] * B.df[2][k] / (l[0] * l[1] * l[2]);
        if (meanold != 0.0)
          for (double &x : u) x = x / meanold * ubulk;
      }
      copy_field(c, (int)b, SNAC_U, nullptr, u.data());
      copy_field(c, (int)b, SNAC_V, nullptr, v.data());
      copy_field(c, (int)b, SNAC_W, nullptr, w.data());
      copy_field(c, (int)b, SNAC_P, nullptr, p.data());
      // initscal (initflow.f90:190-262)
      for (int is = 0; is < cs.nscal; is++) {
        auto bcs = [&](int q, int side) { return B.bcsca[(size_t)is * 6 + q * 2 + side]; };
        std::vector<double> s(ntot, 0.0), s1d(nz + 2, 0.0);
        double sref = 0.5 * (bcs(2, 0) + bcs(2, 1));
        if (cs.is_sforced[is]) sref = cs.scalf[is];
        const std::string in = cs.iniscal[is].substr(0, 3);
        if (in == "zer" || in == "dhc") {
        } else if (in == "uni") {
          for (int k = 1; k <= nz; k++) s1d[k] = sref;
        } else if (in == "cou") {
          for (int k = 1; k <= nz; k++) {
            s1d[k] = 0.5 * (1.0 - 2.0 * (zc[k] / l[2])) * 1.0 + 0.5;
            s1d[k] = bcs(2, 0) * s1d[k] + bcs(2, 1) * (1.0 - s1d[k]);
          }
        } else {
          throw std::runtime_error("ERROR: invalid name for initial scalar field");
        }
        for (int k = 1; k <= nz; k++)
          for (int j = 1; j <= ny; j++)
            for (int i = 1; i <= nx; i++) s[at(i, j, k)] = s1d[k];
        if (in == "dhc")
          for (int k = 1; k <= nz; k++)
            for (int j = 1; j <= ny; j++)
              for (int i = 1; i <= nx; i++) {
                const double xx = xc[i] / l[0];
                s[at(i, j, k)] = (1.0 - xx) * bcs(0, 0) + xx * bcs(0, 1);
              }
        copy_field(c, (int)b, SNAC_SCAL0 + 2 * is, nullptr, s.data());
      }
    }
  });
}

int snac_ctx_get_run_params(const snac_ctx *c, int *iparams) {
  return guard([&] {
    const Case &cs = checked(const_cast<snac_ctx *>(c))->cs;
    const int v[8] = {cs.iout0d, cs.iout1d, cs.iout2d, cs.iout3d, cs.isave, cs.nsaves_max, cs.is_overwrite_save ? 1 : 0,
                      cs.restart ? 1 : 0};
    std::copy(v, v + 8, iparams);
  });
}

int snac_save_grid(const snac_ctx *c, const char *datadir) {
  return guard([&] {
    const Case &cs = checked(const_cast<snac_ctx *>(c))->cs;
    const std::string dir = datadir ? datadir : "";
    for (size_t b = 0; b < cs.blocks.size(); b++) {
      const Block &B = cs.blocks[b];
      int lo[3], hi[3];
      double lmin[3], lmax[3];
      for (int q = 0; q < 3; q++) {
        save_grid(dir + "grid_" + "xyz"[q] + "_b_" + block_num((int)b), B.n[q], B.rf[q], B.rc[q], B.df0[q], B.dc0[q]);
        lo[q] = B.lo_g[q];
        hi[q] = B.lo_g[q] + B.n[q] - 1;
        lmin[q] = B.lmin[q];
        lmax[q] = B.lmax[q];
      }
      write_geometry(dir + "geometry_b_" + block_num((int)b) + ".out", lo, hi, lmin, lmax);
    }
  });
}

int snac_rk_mom(snac_ctx *c, int irk, double dt) {
  return guard([&] {
    on_device(c);
    if (irk < 1 || irk > 3) throw std::runtime_error("irk must be 1..3");
    // standalone call: carry the current ghost layer over, as the in-place
    // reference update leaves it untouched
    int nx = 1 - c->cur;
    for (auto &D : c->blk)
      for (int q = 0; q < 3; q++)
        SNAC_HIP(hipMemcpyAsync(D.vel[nx][q], D.vel[c->cur][q], D.ntot * sizeof(double), hipMemcpyDeviceToDevice,
                                c->stream));
    rk_mom(c, irk, dt);
  });
}

int snac_bounduvw(snac_ctx *c, int is_correc) {
  return guard([&] { run_program(on_device(c), c->prog_bound_only[c->cur][is_correc ? 1 : 0]); });
}

int snac_inflow(snac_ctx *c, int is_correc) {
  return guard([&] { run_program(on_device(c), c->prog_inflow[c->cur][is_correc ? 1 : 0]); });
}

int snac_boundp(snac_ctx *c, int fid) {
  return guard([&] {
    on_device(c);
    if (fid == SNAC_P) run_program(c, c->prog_p);
    else if (fid == SNAC_PP) run_program(c, c->prog_pp);
    else if (fid >= SNAC_SCAL0 && fid < SNAC_SCAL0 + 2 * c->cs.nscal && (fid - SNAC_SCAL0) % 2 == 0)
      run_program(c, c->prog_scal_one[c->scur][(fid - SNAC_SCAL0) / 2]);
    else throw std::runtime_error("boundp: field must be SNAC_P, SNAC_PP or a scalar");
  });
}

int snac_fillps(snac_ctx *c, double dtrk) { return guard([&] { fillps(on_device(c), dtrk); }); }

int snac_solve_pressure(snac_ctx *c) { return guard([&] { solve_pressure(on_device(c)); }); }

int snac_correc(snac_ctx *c, double dtrk) { return guard([&] { correc(on_device(c), dtrk, false); }); }

int snac_updt_pressure(snac_ctx *c) {
  return guard([&] {
    for (auto &D : on_device(c)->blk) launch_updt_pressure(c->stream, D.g, D.pp, D.p);
  });
}

int snac_updt_pressure_impdiff(snac_ctx *c, double alpha) {
  return guard([&] {
    for (auto &D : on_device(c)->blk) launch_updt_pressure_imp(c->stream, D.g, D.m, D.pp, D.p, alpha);
  });
}

int snac_set_buoyancy(snac_ctx *c, int enable) {
  return guard([&] {
    checked(c);
    if (enable && c->cs.nscal <= 0) throw std::runtime_error("_BOUSSINESQ_BUOYANCY requires nscal > 0.");
    if (c->stream) SNAC_HIP(hipStreamSynchronize(c->stream));
    c->buoyancy = enable != 0;
    drop_graph(c);
  });
}

int snac_set_impdiff(snac_ctx *c, int enable) {
  return guard([&] {
    checked(c);  // a host-only context plans (and validates) without device arrays
    if (c->stream) SNAC_HIP(hipStreamSynchronize(c->stream));
    c->impdiff = false;
    if (enable) {
      plan_impdiff(c);
      c->impdiff = true;
    }
    drop_graph(c);
  });
}

// the implicit-diffusion field of a field id: 0..2 = u, v, w, 3+i = scalar i
int impdiff_comp(snac_ctx *c, int fid, const char *what) {
  if (!c->impdiff) throw std::runtime_error("implicit diffusion is off (snac_set_impdiff)");
  int comp = -1;
  if (fid >= SNAC_U && fid <= SNAC_W) comp = fid - SNAC_U;
  else if (fid >= SNAC_SCAL0 && fid < SNAC_SCAL0 + 2 * c->cs.nscal && (fid - SNAC_SCAL0) % 2 == 0)
    comp = 3 + (fid - SNAC_SCAL0) / 2;
  if (comp < 0) throw std::runtime_error(std::string(what) + ": field must be u, v, w or a scalar");
  return comp;
}

int snac_impdiff_stats(snac_ctx *c, int fid, int *iters, double *max_relres) {
  return guard([&] {
    on_device(c);
    const ImpField &f = c->imp[impdiff_comp(c, fid, "impdiff_stats")];
    int it = 0;
    double rel = 0.0;
    if (f.mg && f.mg->ok) mg_stats(*f.mg, c->stream, it, rel);
    if (iters) *iters = it;
    if (max_relres) *max_relres = rel;
  });
}

int snac_solve_impdiff(snac_ctx *c, int fid, double alpha) {
  return guard([&] {
    on_device(c);
    const int comp = impdiff_comp(c, fid, "solve_impdiff");
    if (!(alpha < 0.0)) throw std::runtime_error("solve_impdiff: alpha must be negative (-visc*dtrk/2)");
    solve_impdiff(c, comp, alpha);
    std::fill(c->tcnt, c->tcnt + SNAC_NPHASE, 0);
  });
}

int snac_chkdiv(snac_ctx *c, double *divtot, double *divmax) {
  return guard([&] {
    double r[2];
    reduce(on_device(c), true, r);
    if (divtot) *divtot = r[0];
    if (divmax) *divmax = r[1];
  });
}

int snac_chkdt(snac_ctx *c, double *dtmax) {
  return guard([&] {
    double r[2];
    reduce(on_device(c), false, r);
    double dti = r[1] == 0.0 ? 1.0 : r[1];
    double v = std::sqrt(3.0) / dti;
    // main.f90:660 passes max(visc, alpha_max) as the diffusivity; the
    // diffusive limit does not apply with implicit diffusion (chkdt.f90:51)
    const double ve = c->cs.nscal > 0 ? std::max(c->cs.visc, c->cs.alpha_max) : c->cs.visc;
    if (!c->impdiff) v = std::min(v, 1.65 / 12.0 / ve * c->dlmin * c->dlmin);
    if (dtmax) *dtmax = v;
  });
}

int snac_init_bounds(snac_ctx *c) {
  return guard([&] {
    run_program(on_device(c), c->prog_bound_only[c->cur][0]);
    run_program(c, c->prog_p);
    if (c->cs.nscal > 0) run_program(c, c->prog_scal[c->scur]);
    run_program(c, c->prog_inflow[c->cur][0]);
  });
}

int snac_rk_scal(snac_ctx *c, int irk, double dt) {
  return guard([&] {
    on_device(c);
    if (irk < 1 || irk > 3) throw std::runtime_error("irk must be 1..3");
    rk_scal(c, irk, dt);
  });
}

int snac_scalar_forcing(snac_ctx *c, double *f) {
  return guard([&] {
    on_device(c);
    if (c->cs.nscal == 0) return;
    SNAC_HIP(hipMemcpyAsync(f, c->sforce, c->cs.nscal * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    SNAC_HIP(hipStreamSynchronize(c->stream));
  });
}

int snac_step(snac_ctx *c, double dt) { return guard([&] { run_steps(on_device(c), dt, 1); }); }

int snac_steps(snac_ctx *c, double dt, int nsteps) { return guard([&] { run_steps(on_device(c), dt, nsteps); }); }

int snac_set_graph(snac_ctx *c, int enable) {
  return guard([&] {
    on_device(c);
    SNAC_HIP(hipStreamSynchronize(c->stream));
    c->use_graph = enable != 0;
    if (!c->use_graph) drop_graph(c);
  });
}

int snac_set_timing(snac_ctx *c, int enable) {
  return guard([&] {
    on_device(c);
    collect_timing(c);
    c->timing = enable != 0;
    std::fill(c->tacc, c->tacc + SNAC_NPHASE, 0.0);
    std::fill(c->tcnt, c->tcnt + SNAC_NPHASE, 0);
    // the finest MG level's colour sweeps and the fused coarse end get
    // their own events
    auto marker = [c](int phase) {
      auto open = std::make_shared<hipEvent_t>();
      return [c, open, phase](bool begin) {
        if (begin) {
          *open = ev_get(c);
          SNAC_HIP(hipEventRecord(*open, c->stream));
        } else {
          hipEvent_t b = ev_get(c);
          if (hipEventRecord(b, c->stream) == hipSuccess) c->pending.push_back({phase, *open, b});
        }
      };
    };
    if (c->timing) {
      c->mg.mark_line0 = marker(SNAC_PH_MG_LINE0);
      c->mg.mark_tail = marker(SNAC_PH_MG_TAIL);
      c->mg.mark_update = marker(SNAC_PH_MG_UPDATE);
    } else {
      c->mg.mark_line0 = nullptr;
      c->mg.mark_tail = nullptr;
      c->mg.mark_update = nullptr;
    }
  });
}

int snac_get_timing(snac_ctx *c, double *ms, int *count) {
  return guard([&] {
    on_device(c);
    collect_timing(c);
    for (int q = 0; q < SNAC_NPHASE; q++) {
      if (ms) ms[q] = c->tacc[q];
      if (count) count[q] = c->tcnt[q];
    }
    std::fill(c->tacc, c->tacc + SNAC_NPHASE, 0.0);
    std::fill(c->tcnt, c->tcnt + SNAC_NPHASE, 0);
  });
}

int snac_get_autotune(snac_ctx *c, int *choices, double *ms, int n) {
  return guard([&] {
    on_device(c);
    const TensorPlan &t = c->tp;
    std::vector<int> ch;
    for (auto &p : t.P.fwd) ch.push_back(p.fft ? p.lshift : -1);
    for (auto &p : t.P.bwd) ch.push_back(p.fft ? p.lshift : -1);
    std::vector<double> tm = t.tune_ms;
    if (t.ok) ch.push_back(t.tvariant);
    ch.push_back(c->rk_variant);
    tm.push_back(c->rk_tune_ms);
    ch.push_back(c->rk32_variant);
    tm.push_back(c->rk32_tune_ms);
    for (int q = 0; q < n; q++) {
      if (choices) choices[q] = q < (int)ch.size() ? ch[q] : -2;
      if (ms) ms[q] = q < (int)tm.size() ? tm[q] : 0.0;
    }
  });
}

int snac_mg_line_choices(snac_ctx *c, int *v64, int *v32, int n) {
  return guard([&] {
    const MGPlan &P = c->mg;
    for (int l = 0; l < n; l++) {
      if (v64) v64[l] = c->mg_planned && l < (int)P.h64.lev.size() ? P.h64.lev[l].lvar : -1;
      if (v32) v32[l] = c->mg_planned && l < (int)P.h32.lev.size() ? P.h32.lev[l].lvar : -1;
    }
  });
}

int snac_set_kernel_variant(snac_ctx *c, int kernel, int variant) {
  return guard([&] {
    on_device(c);
    if (kernel == SNAC_KV_RK_MOM) {
      if (variant < 0 || variant > 16) throw std::runtime_error("rk_mom variant must be 0..16");
      c->rk_variant = variant;
    } else if (kernel == SNAC_KV_MG_LINE) {
      if (variant > 4) throw std::runtime_error("MG line relaxation variant must be 0..4 or negative");
      if (c->stream) SNAC_HIP(hipStreamSynchronize(c->stream));
      mg_set_line_variant(variant < 0 ? -1 : variant);
    } else if (kernel == SNAC_KV_MG_TAIL) {
      if (variant < 0 || variant > 1) throw std::runtime_error("MG tail variant must be 0 or 1");
      if (c->stream) SNAC_HIP(hipStreamSynchronize(c->stream));
      mg_set_tail_lds(variant == 1);
    } else if (kernel == SNAC_KV_MG_CG) {
      if (variant < 0 || variant > 1) throw std::runtime_error("MG CG direction-pass variant must be 0 or 1");
      if (c->stream) SNAC_HIP(hipStreamSynchronize(c->stream));
      drop_graph(c);
      mg_set_cg_lds(variant == 1);
    } else if (kernel == SNAC_KV_MG_WARM) {
      if (variant < 0 || variant > 1) throw std::runtime_error("MG warm-start variant must be 0 or 1");
      drop_graph(c);
      c->mg_warm = variant == 1;
      c->mg.warm = c->mg_warm;
      c->mg.fixed_iters = 0;  // the graph's CG count is measured again
      c->mg.max_seen = 0;
    } else if (kernel == SNAC_KV_MG_RCP) {
      if (variant < 0 || variant > 1) throw std::runtime_error("MG reciprocal variant must be 0 or 1");
      if (c->stream) SNAC_HIP(hipStreamSynchronize(c->stream));
      mg_set_fast_rcp(variant == 1);
    } else if (kernel == SNAC_KV_RK_MOM32) {
      if (variant < 7 || variant > 16) throw std::runtime_error("FP32 rk_mom variant must be 7..16");
      c->rk32_variant = variant;
    } else if (kernel == SNAC_KV_TRIDIAG) {
      if (!c->tp.ok) throw std::runtime_error("no tensor-product plan on this case");
      const int m = c->tp.G[c->tp.s];
      if (variant < 0 || variant > 3 || (variant == 0 && m > 128) || (variant == 2 && tridiag_chunks(m) == 0) ||
          (variant == 3 && tridiag_chunks(m, true) == 0))
        throw std::runtime_error("tridiagonal variant not available for lines of " + std::to_string(m));
      c->tp.tvariant = variant;
    } else if (kernel == SNAC_KV_GHOST) {
      if (variant < 0 || variant > 1) throw std::runtime_error("ghost-fill variant must be 0 or 1");
      c->ghost_fused = variant == 1;
    } else {
      throw std::runtime_error("unknown kernel id");
    }
    drop_graph(c);
  });
}

int snac_sync(snac_ctx *c) { return guard([&] { SNAC_HIP(hipStreamSynchronize(on_device(c)->stream)); }); }

int snac_r2r_lines(snac_ctx *c, int kind, int n, int howmany, int strided, int tile, const double *in,
                   double *out) {
  return guard([&] {
    on_device(c);
    if (tile < 0 || tile > 7) throw std::runtime_error("r2r_lines: tile must be 0..7");
    R2RPlan pl = r2r_plan(kind, n);
    pl.lshift = tile;
    size_t tot = (size_t)n * howmany;
    double *din = nullptr, *dout = nullptr;
    SNAC_HIP(hipMalloc(&din, tot * sizeof(double)));
    SNAC_HIP(hipMalloc(&dout, tot * sizeof(double)));
    SNAC_HIP(hipMemcpyAsync(din, in, tot * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if (strided) launch_r2r(c->stream, pl, din, howmany, 1, 0, dout, howmany, 1, 0, howmany, 1, 1.0);
    else launch_r2r(c->stream, pl, din, 1, n, 0, dout, 1, n, 0, howmany, 1, 1.0);
    SNAC_HIP(hipMemcpyAsync(out, dout, tot * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    SNAC_HIP(hipStreamSynchronize(c->stream));
    SNAC_HIP(hipFree(din));
    SNAC_HIP(hipFree(dout));
  });
}

// ---- the learned preconditioner's network on its own (training, checks)
struct snac_net {
  snac::Net *net = nullptr;
  hipStream_t stream = nullptr;
  float *x = nullptr, *t = nullptr, *y = nullptr, *loss = nullptr;
  size_t cap = 0;  // samples the staging buffers hold
};

namespace {
void net_stage(snac_net *n, size_t samples) {
  if (samples <= n->cap) return;
  for (float *p : {n->x, n->t, n->y, n->loss}) if (p) (void)hipFree(p);
  const size_t pl = (size_t)n->net->H() * n->net->W();
  SNAC_HIP(hipMalloc(&n->x, samples * pl * 2 * sizeof(float)));
  SNAC_HIP(hipMalloc(&n->t, samples * pl * sizeof(float)));
  SNAC_HIP(hipMalloc(&n->y, samples * pl * sizeof(float)));
  SNAC_HIP(hipMalloc(&n->loss, sizeof(float)));
  n->cap = samples;
}
snac_net *net_checked(snac_net *n, int B = 1) {
  if (!n || !n->net) throw std::runtime_error("null network");
  if (B < 1 || B > n->net->maxbatch())
    throw std::runtime_error("network batch " + std::to_string(B) + " outside 1.." + std::to_string(n->net->maxbatch()));
  return n;
}
}  // namespace

int snac_net_create(snac_net **out, int H, int W, int maxbatch, unsigned seed, int device) {
  return guard([&] {
    if (!out) throw std::runtime_error("null argument");
    *out = nullptr;
    SNAC_HIP(hipSetDevice(device));
    auto *n = new snac_net();
    try {
      n->net = new Net(H, W, maxbatch, seed);
      SNAC_HIP(hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking));
    } catch (...) {
      delete n->net;
      delete n;
      throw;
    }
    *out = n;
  });
}

int snac_net_destroy(snac_net *n) {
  if (!n) return 0;
  if (n->stream) (void)hipStreamSynchronize(n->stream);
  for (float *p : {n->x, n->t, n->y, n->loss}) if (p) (void)hipFree(p);
  delete n->net;
  if (n->stream) (void)hipStreamDestroy(n->stream);
  delete n;
  return 0;
}

int snac_net_conv_variant(int variant) { return guard([&] { net_set_conv_variant(variant); }); }

long snac_net_num_params(const snac_net *n) { return n && n->net ? (long)n->net->num_params() : -1; }

int snac_net_get_params(const snac_net *n, float *params) {
  return guard([&] { net_checked(const_cast<snac_net *>(n))->net->get_params(params); });
}

int snac_net_set_params(snac_net *n, const float *params) {
  return guard([&] { net_checked(n)->net->set_params(params); });
}

int snac_net_forward(snac_net *n, int B, const float *x, float *y) {
  return guard([&] {
    net_checked(n, B);
    net_stage(n, B);
    const size_t pl = (size_t)n->net->H() * n->net->W();
    SNAC_HIP(hipMemcpyAsync(n->x, x, (size_t)B * pl * 2 * sizeof(float), hipMemcpyHostToDevice, n->stream));
    n->net->set_input2(n->stream, B, n->x);
    n->net->forward(n->stream, B);
    n->net->get_output(n->stream, B, n->y);
    SNAC_HIP(hipMemcpyAsync(y, n->y, (size_t)B * pl * sizeof(float), hipMemcpyDeviceToHost, n->stream));
    SNAC_HIP(hipStreamSynchronize(n->stream));
  });
}

int snac_net_loss_grad(snac_net *n, int B, const float *x, const float *t, float *loss, float *grad) {
  return guard([&] {
    net_checked(n, B);
    net_stage(n, B);
    const size_t pl = (size_t)n->net->H() * n->net->W();
    SNAC_HIP(hipMemcpyAsync(n->x, x, (size_t)B * pl * 2 * sizeof(float), hipMemcpyHostToDevice, n->stream));
    SNAC_HIP(hipMemcpyAsync(n->t, t, (size_t)B * pl * sizeof(float), hipMemcpyHostToDevice, n->stream));
    n->net->set_input2(n->stream, B, n->x);
    n->net->forward(n->stream, B);
    n->net->loss_backward(n->stream, B, n->t, n->loss);
    SNAC_HIP(hipMemcpyAsync(loss, n->loss, sizeof(float), hipMemcpyDeviceToHost, n->stream));
    SNAC_HIP(hipStreamSynchronize(n->stream));
    if (grad) n->net->grads_to_host(grad);
  });
}

int snac_net_train(snac_net *n, int nsamples, const float *x, const float *t, int batch, int epochs, float lr,
                   float *loss_per_epoch) {
  return guard([&] {
    net_checked(n);
    if (nsamples < 1 || batch < 1 || batch > n->net->maxbatch() || epochs < 1)
      throw std::runtime_error("net_train: invalid sizes");
    net_stage(n, nsamples);
    const size_t pl = (size_t)n->net->H() * n->net->W();
    SNAC_HIP(hipMemcpyAsync(n->x, x, (size_t)nsamples * pl * 2 * sizeof(float), hipMemcpyHostToDevice, n->stream));
    SNAC_HIP(hipMemcpyAsync(n->t, t, (size_t)nsamples * pl * sizeof(float), hipMemcpyHostToDevice, n->stream));
    for (int e = 0; e < epochs; e++) {
      double sum = 0.0;
      int nb = 0;
      for (int s0 = 0; s0 + batch <= nsamples; s0 += batch, nb++) {
        n->net->set_input2(n->stream, batch, n->x + (size_t)s0 * pl * 2);
        n->net->forward(n->stream, batch);
        n->net->loss_backward(n->stream, batch, n->t + (size_t)s0 * pl, n->loss);
        n->net->adam_step(n->stream, lr);
        float l = 0.f;
        SNAC_HIP(hipMemcpyAsync(&l, n->loss, sizeof(float), hipMemcpyDeviceToHost, n->stream));
        SNAC_HIP(hipStreamSynchronize(n->stream));
        sum += l;
      }
      if (loss_per_epoch) loss_per_epoch[e] = (float)(sum / std::max(nb, 1));
    }
  });
}

int snac_stream_copy(snac_ctx *c, long n, int reps, double *gbps) {
  return guard([&] {
    on_device(c);
    if (n < 2 || reps < 1) throw std::runtime_error("stream_copy: n >= 2 and reps >= 1 required");
    n &= ~1L;
    double *a = nullptr, *b = nullptr;
    SNAC_HIP(hipMalloc(&a, n * sizeof(double)));
    SNAC_HIP(hipMalloc(&b, n * sizeof(double)));
    SNAC_HIP(hipMemsetAsync(a, 0, n * sizeof(double), c->stream));
    launch_stream_copy(c->stream, a, b, n);  // warm-up
    hipEvent_t e0, e1;
    SNAC_HIP(hipEventCreate(&e0));
    SNAC_HIP(hipEventCreate(&e1));
    SNAC_HIP(hipEventRecord(e0, c->stream));
    for (int r = 0; r < reps; r++) launch_stream_copy(c->stream, a, b, n);
    SNAC_HIP(hipEventRecord(e1, c->stream));
    SNAC_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    SNAC_HIP(hipEventElapsedTime(&ms, e0, e1));
    if (gbps) *gbps = 16.0 * (double)n * reps / (ms * 1e-3) / 1e9;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    SNAC_HIP(hipFree(a));
    SNAC_HIP(hipFree(b));
  });
}

}  // extern "C"
