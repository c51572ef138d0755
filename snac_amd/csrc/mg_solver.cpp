// Host side of the batched MG-preconditioned CG (mg.hpp): operator assembly
// on the bounding box of the union (the formulas of init_matrix_2d
// solver.f90:687-842 with lambda*dxf*dyf on the diagonal, init_n_2d_matrices
// 981-1020, and of init_matrix_3d 158-372), coarse levels by semi-coarsening
// along a (and m on the 3D path), the V-cycle and the CG iteration of all
// groups at once.
//
// Coarse operators: cells are merged in pairs along the coarsened axis and
// the operator is re-discretised on the merged cells rather than formed as
// P^T A P with piecewise-constant P.  The Galerkin product keeps the fine
// coupling across a merged face although the coarse cell centres are twice
// as far apart, so every level would double the coarsened-direction
// coupling relative to the others and the V-cycle would degrade with depth.
// Re-discretised: couplings across the line axis add up (the merged face is
// the sum of the faces), couplings along a coarsened axis and the Dirichlet
// boundary terms in that direction are halved (twice the distance), the
// lambda weight adds up (the merged cell area), and the diagonal is the sum
// of the couplings plus the boundary terms, so Neumann/periodic problems
// stay singular with the constant null vector.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <sstream>
#include <string>

#include "mg.hpp"
#include "net.hpp"

namespace snac {
namespace {

template <class T>
T *dev_upload(MGPlan &P, const std::vector<T> &h) {
  void *p = nullptr;
  SNAC_HIP(hipMalloc(&p, std::max<size_t>(h.size(), 1) * sizeof(T)));
  if (!h.empty()) SNAC_HIP(hipMemcpy(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  P.owned.push_back(p);
  return (T *)p;
}

template <class T>
T *dev_zeros(MGPlan &P, size_t n) {
  void *p = nullptr;
  SNAC_HIP(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
  SNAC_HIP(hipMemset(p, 0, std::max<size_t>(n, 1) * sizeof(T)));
  P.owned.push_back(p);
  return (T *)p;
}

// host copy of one level's coefficients, dims (nmq, na, ns), nmq = nm when
// coupled, else 1
struct HostLev {
  int nm, na, ns, nmq;
  bool per_a, per_m;
  std::vector<double> dg, w, cs, ca, cm;
  std::vector<double> bd[3];  // boundary terms of the diagonal from the s / a / m faces
  std::vector<unsigned char> act;
  long at(int m, int a, int s) const { return ((long)m * na + a) * ns + s; }
  void resize() {
    size_t n = (size_t)nmq * na * ns;
    dg.assign(n, 1.0); w.assign(n, 0.0); cs.assign(n, 0.0); ca.assign(n, 0.0); cm.assign(n, 0.0);
    for (auto &v : bd) v.assign(n, 0.0);
    act.assign(n, 0);
  }
  // diagonal = couplings to all neighbours + boundary terms (active cells)
  void set_diagonal() {
    for (int m = 0; m < nmq; m++)
      for (int a = 0; a < na; a++)
        for (int s = 0; s < ns; s++) {
          const long q = at(m, a, s);
          if (!act[q]) { dg[q] = 1.0; continue; }
          double d = bd[0][q] + bd[1][q] + bd[2][q] + cs[q] + (s > 0 ? cs[q - 1] : 0.0);
          d += ca[q] + (a > 0 ? ca[at(m, a - 1, s)] : (per_a ? ca[at(m, na - 1, s)] : 0.0));
          d += cm[q] + (m > 0 ? cm[at(m - 1, a, s)] : (per_m ? cm[at(nmq - 1, a, s)] : 0.0));
          dg[q] = d;
        }
  }
};

bool uniform_axis(const Case &cs, int a) {
  double h = cs.blocks[0].df[a][1];
  for (auto &B : cs.blocks)
    for (int i = 0; i <= B.n[a] + 1; i++) {
      if (std::fabs(B.df[a][i] - h) > 1e-12 * h) return false;
      if (i <= B.n[a] && std::fabs(B.dc[a][i] - h) > 1e-12 * h) return false;
    }
  return true;
}

HostLev coarsen(const HostLev &F, bool coupled) {
  HostLev C;
  const bool ra = F.na > 1, rm = coupled && F.nm > 1;
  C.na = ra ? (F.na + 1) / 2 : F.na;
  C.nm = rm ? (F.nm + 1) / 2 : F.nm;
  C.ns = F.ns;
  C.nmq = coupled ? C.nm : 1;
  C.per_a = F.per_a && C.na > 1;
  C.per_m = F.per_m && C.nm > 1;
  C.resize();
  const double fa = ra ? 0.5 : 1.0, fm = rm ? 0.5 : 1.0;
  for (int M = 0; M < C.nmq; M++)
    for (int I = 0; I < C.na; I++)
      for (int s = 0; s < C.ns; s++) {
        const long qc = C.at(M, I, s);
        const int m0 = rm ? 2 * M : M, m1 = rm ? std::min(2 * M + 1, F.nmq - 1) : M;
        const int a0 = ra ? 2 * I : I, a1 = ra ? std::min(2 * I + 1, F.na - 1) : I;
        bool act = false;
        double w = 0.0, csum = 0.0, bds = 0.0, bda = 0.0, bdm = 0.0, ca = 0.0, cm = 0.0;
        for (int m = m0; m <= m1; m++)
          for (int a = a0; a <= a1; a++) {
            const long q = F.at(m, a, s);
            csum += F.cs[q];
            if (!F.act[q]) continue;
            act = true;
            w += F.w[q];
            bds += F.bd[0][q];
            bda += F.bd[1][q];
            bdm += F.bd[2][q];
          }
        // the link leaving the merged cell towards +a / +m (internal when the
        // coarse axis collapses to one periodic cell)
        if (C.na > 1 || !ra)
          for (int m = m0; m <= m1; m++) ca += F.ca[F.at(m, a1, s)];
        if (coupled && (C.nm > 1 || !rm))
          for (int a = a0; a <= a1; a++) cm += F.cm[F.at(m1, a, s)];
        if (!act) continue;  // inactive: identity row
        C.act[qc] = 1;
        C.w[qc] = w;
        C.cs[qc] = csum;
        C.ca[qc] = fa * ca;
        C.cm[qc] = fm * cm;
        C.bd[0][qc] = bds;
        C.bd[1][qc] = fa * bda;
        C.bd[2][qc] = fm * bdm;
      }
  C.set_diagonal();
  return C;
}

}  // namespace

void mg_free(MGPlan &P) {
  delete P.net;
  P.net = nullptr;
  for (void *p : P.owned) (void)hipFree(p);
  P.owned.clear();
  if (P.left_host) (void)hipHostFree(P.left_host);
  P.left_host = nullptr;
  if (P.stall_host) (void)hipHostFree(P.stall_host);
  P.stall_host = nullptr;
}

void mg_plan(MGPlan &P, const Case &cs, int device, int comp) {
  P.ok = false;
  P.comp = comp;
  const int fa = cs.fft_axis - 1;
  // the field: the pressure (comp -1), or the Helmholtz operator of a
  // velocity component (staggered along its own axis) or a scalar
  const int stag = comp >= 0 && comp < 3 ? comp : -1;
  auto cbc = [&](const Block &B, int a, int sd) -> char {
    if (comp < 0) return B.cbcpre[a][sd];
    if (comp < 3) return B.cbcvel[comp][a][sd];
    return B.cbcsca[((size_t)(comp - 3) * 3 + a) * 2 + sd];
  };
  // unknowns of a block along an axis: a staggered axis bounded at its upper
  // end keeps the boundary face out (main.f90:774-775)
  auto extent = [&](const Block &B, int a) { return B.n[a] - ((a == stag && B.is_bound[a][1]) ? 1 : 0); };
  // spacing to the lower / upper neighbour of unknown i and its width dx2
  auto sp_lo = [&](const Block &B, int a, int i) { return a == stag ? B.df[a][i] : B.dc[a][i - 1]; };
  auto sp_hi = [&](const Block &B, int a, int i) { return a == stag ? B.df[a][i + 1] : B.dc[a][i]; };
  auto wdx2 = [&](const Block &B, int a, int i) { return a == stag ? B.dc[a][i] : B.df[a][i]; };
  P.fa = fa;
  const bool coupled = fa < 0;
  int gmax[3];
  for (int a = 0; a < 3; a++) {
    P.gmin[a] = 1 << 30;
    gmax[a] = -(1 << 30);
    for (auto &B : cs.blocks) {
      P.gmin[a] = std::min(P.gmin[a], B.lo_g[a]);
      gmax[a] = std::max(gmax[a], B.lo_g[a] + B.n[a] - 1);
    }
    P.G[a] = gmax[a] - P.gmin[a] + 1;
  }
  if (fa >= 0)
    for (auto &B : cs.blocks)
      if (B.n[fa] != P.G[fa]) { P.why = "a block does not span the whole FFT axis"; return; }
  if (comp >= 0 && fa < 0) { P.why = "implicit diffusion on the MG-PCG path needs an FFT axis"; return; }
  // a component staggered along a bounded FFT axis keeps its upper boundary
  // face out: the transform covers the first n-1 planes (fftini
  // fft.f90:31-33), n-1 modes
  const bool fa_centred = fa >= 0 && fa != stag;
  const int qfft = (fa >= 0 && !fa_centred && cbc(cs.blocks[0], fa, 0) != 'F') ? 1 : 0;
  // axes: s = line axis (non-periodic; stretched preferred, then longest)
  std::vector<int> cand;
  for (int a = 0; a < 3; a++)
    if (a != fa) cand.push_back(a);
  int s = -1;
  for (int pass = 0; pass < 2 && s < 0; pass++)
    for (int a : cand) {
      if (cs.periods[a] != 0) continue;
      if (pass == 0 && uniform_axis(cs, a)) continue;
      if (s < 0 || P.G[a] > P.G[s]) s = a;
    }
  if (s < 0) { P.why = "every axis of the Poisson problem is periodic"; return; }
  std::vector<int> rest;
  for (int a : cand)
    if (a != s) rest.push_back(a);
  P.ax_s = s;
  if (coupled) {
    // a: the larger of the remaining axes; m: the other one
    if (P.G[rest[0]] >= P.G[rest[1]]) { P.ax_a = rest[0]; P.ax_m = rest[1]; }
    else { P.ax_a = rest[1]; P.ax_m = rest[0]; }
  } else {
    P.ax_a = rest[0];
    P.ax_m = fa;
  }
  const int ns = P.G[P.ax_s], na = P.G[P.ax_a], nm = P.G[P.ax_m] - (coupled ? 0 : qfft);
  if (ns > 64 * 17) { P.why = "line axis longer than 1088 cells"; return; }

  // fine-level operator
  HostLev L0;
  L0.nm = nm; L0.na = na; L0.ns = ns; L0.nmq = coupled ? nm : 1;
  L0.per_a = cs.periods[P.ax_a] != 0;
  L0.per_m = coupled && cs.periods[P.ax_m] != 0;
  L0.resize();
  std::vector<double> cmm((size_t)L0.nmq * na * ns * 3, 0.0);  // minus-side couplings (check only)
  const int nd = coupled ? 3 : 2;
  int ax[3];
  if (coupled) { ax[0] = P.ax_s; ax[1] = P.ax_a; ax[2] = P.ax_m; }
  else { ax[0] = P.ax_s; ax[1] = P.ax_a; ax[2] = -1; }
  bool has_dirichlet = false;
  for (auto &B : cs.blocks) {
    const int n0 = extent(B, ax[0]), n1 = extent(B, ax[1]), n2 = coupled ? extent(B, ax[2]) : 1;
    const int nn[3] = {n0, n1, n2};
    for (int i2 = 1; i2 <= n2; i2++)
      for (int i1 = 1; i1 <= n1; i1++)
        for (int i0 = 1; i0 <= n0; i0++) {
          const int idx[3] = {i0, i1, i2};
          double weight = 1.0;
          for (int e = 0; e < nd; e++) weight *= wdx2(B, ax[e], idx[e]);
          double cmv[3], cpv[3], ssum = 0.0;
          for (int e = 0; e < nd; e++) {
            double other = 1.0;
            for (int o = 0; o < nd; o++)
              if (o != e) other *= wdx2(B, ax[o], idx[o]);
            cmv[e] = other / sp_lo(B, ax[e], idx[e]);
            cpv[e] = other / sp_hi(B, ax[e], idx[e]);
            ssum += cmv[e] + cpv[e];
          }
          double diag0 = -ssum;
          double bdv[3] = {0.0, 0.0, 0.0};  // (1 - sgn) * c of the boundary faces
          for (int e = 0; e < nd; e++) {
            const int a = ax[e];
            if (cs.periods[a] != 0) continue;
            // a Dirichlet value lies on the boundary face itself for a
            // staggered axis (sgn 0), half a cell out otherwise (sgn -1)
            const double sgd = a == stag ? 0.0 : -1.0;
            if (B.is_bound[a][0] && idx[e] == 1) {
              const char bc = cbc(B, a, 0);
              const double sg = bc == 'N' ? 1.0 : (bc == 'D' ? sgd : 0.0);
              if (bc == 'D') has_dirichlet = true;
              diag0 += sg * cmv[e];
              bdv[e] += (1.0 - sg) * cmv[e];
              cmv[e] = 0.0;
            }
            if (B.is_bound[a][1] && idx[e] == nn[e]) {
              const char bc = cbc(B, a, 1);
              const double sg = bc == 'N' ? 1.0 : (bc == 'D' ? sgd : 0.0);
              if (bc == 'D') has_dirichlet = true;
              diag0 += sg * cpv[e];
              bdv[e] += (1.0 - sg) * cpv[e];
              cpv[e] = 0.0;
            }
          }
          const int gs = B.lo_g[ax[0]] - P.gmin[ax[0]] + i0 - 1;
          const int ga = B.lo_g[ax[1]] - P.gmin[ax[1]] + i1 - 1;
          const int gm = coupled ? B.lo_g[ax[2]] - P.gmin[ax[2]] + i2 - 1 : 0;
          const long q = L0.at(gm, ga, gs);
          if (L0.act[q]) { P.why = "overlapping blocks"; return; }
          L0.act[q] = 1;
          L0.dg[q] = -diag0;
          for (int e = 0; e < nd; e++) L0.bd[e][q] = bdv[e];
          L0.w[q] = coupled ? 0.0 : weight;
          L0.cs[q] = cpv[0];
          L0.ca[q] = cpv[1];
          if (coupled) L0.cm[q] = cpv[2];
          for (int e = 0; e < nd; e++) cmm[q * 3 + e] = cmv[e];
        }
  }
  // a Helmholtz operator's unknowns exclude the bounded upper faces of the
  // staggered axis: a link from a neighbouring block into such a face (v at
  // the top of a lower block, whose x-neighbour is the wall face of the block
  // beside it) leaves the union of the blocks' extents, the struct grid of
  // solver.f90:221-224, and meets a zero value there -- it drops out
  if (comp >= 0) {
    double *cpw[3] = {L0.cs.data(), L0.ca.data(), L0.cm.data()};
    const int ext[3] = {ns, na, L0.nmq};
    const bool per[3] = {false, L0.per_a, L0.per_m};
    for (int m = 0; m < L0.nmq; m++)
      for (int a = 0; a < na; a++)
        for (int s2 = 0; s2 < ns; s2++) {
          const long q = L0.at(m, a, s2);
          if (!L0.act[q]) continue;
          const int id[3] = {s2, a, m};
          for (int e = 0; e < nd; e++)
            for (int sd = -1; sd <= 1; sd += 2) {
              int j[3] = {id[0], id[1], id[2]};
              j[e] += sd;
              if (j[e] < 0 || j[e] == ext[e]) {
                if (!per[e]) continue;
                j[e] = j[e] < 0 ? ext[e] - 1 : 0;
              }
              if (L0.act[L0.at(j[2], j[1], j[0])]) continue;
              if (sd > 0) cpw[e][q] = 0.0;
              else cmm[q * 3 + e] = 0.0;
            }
        }
  }
  // every link must be symmetric and join two active cells that are
  // neighbours in the global index space (block faces conform)
  {
    const double *cp[3] = {L0.cs.data(), L0.ca.data(), L0.cm.data()};
    const int ext[3] = {ns, na, L0.nmq};
    const bool per[3] = {false, L0.per_a, L0.per_m};
    for (int m = 0; m < L0.nmq; m++)
      for (int a = 0; a < na; a++)
        for (int s2 = 0; s2 < ns; s2++) {
          const long q = L0.at(m, a, s2);
          if (!L0.act[q]) continue;
          const int id[3] = {s2, a, m};
          for (int e = 0; e < nd; e++) {
            // the + link of this cell against the - link of its neighbour
            int j[3] = {id[0], id[1], id[2]};
            j[e] += 1;
            bool wrap = false;
            if (j[e] == ext[e]) {
              if (!per[e]) {
                if (cp[e][q] != 0.0) { P.why = "internal: coupling across the box boundary"; return; }
                continue;
              }
              j[e] = 0;
              wrap = true;
            }
            (void)wrap;
            const long qn = L0.at(j[2], j[1], j[0]);
            const double c1 = cp[e][q], c2 = L0.act[qn] ? cmm[qn * 3 + e] : 0.0;
            if (std::fabs(c1 - c2) > 1e-12 * std::max(std::fabs(c1), std::fabs(c2))) {
              P.why = "non-conforming block interfaces (asymmetric Poisson operator)";
              return;
            }
          }
          // a - link must lead to an active cell
          for (int e = 0; e < nd; e++) {
            if (cmm[q * 3 + e] == 0.0) continue;
            int j[3] = {id[0], id[1], id[2]};
            j[e] -= 1;
            if (j[e] < 0) {
              if (!per[e]) { P.why = "internal: coupling across the box boundary"; return; }
              j[e] = ext[e] - 1;
            }
            if (!L0.act[L0.at(j[2], j[1], j[0])]) {
              P.why = "non-conforming block interfaces (link to a cell outside the union)";
              return;
            }
          }
        }
  }
  // one-cell periodic axes couple a cell to itself: fold into the diagonal
  if (L0.per_a && na == 1) {
    for (size_t q = 0; q < L0.dg.size(); q++)
      if (L0.act[q]) { L0.dg[q] -= 2.0 * L0.ca[q]; L0.ca[q] = 0.0; }
    L0.per_a = false;
  }
  if (L0.per_m && nm == 1) {
    for (size_t q = 0; q < L0.dg.size(); q++)
      if (L0.act[q]) { L0.dg[q] -= 2.0 * L0.cm[q]; L0.cm[q] = 0.0; }
    L0.per_m = false;
  }
  // a Helmholtz operator's couplings across lines against its weight (the
  // diagonal gains shift * w): how far the shift dominates them
  P.cross_max = 0.0;
  if (comp >= 0 && !coupled)
    for (int a = 0; a < na; a++)
      for (int s2 = 0; s2 < ns; s2++) {
        const long q = L0.at(0, a, s2);
        if (!L0.act[q] || !(L0.w[q] > 0.0)) continue;
        const double cm = a > 0 ? L0.ca[L0.at(0, a - 1, s2)] : (L0.per_a ? L0.ca[L0.at(0, na - 1, s2)] : 0.0);
        P.cross_max = std::max(P.cross_max, (L0.ca[q] + cm) / L0.w[q]);
      }
  // zero row sums at lambda = 0: all-Neumann/periodic, singular
  bool zero_rowsum = !has_dirichlet;
  double nact_plane = 0.0;
  for (size_t q = 0; q < L0.dg.size(); q++) if (L0.act[q]) nact_plane += 1.0;

  // hierarchy
  std::vector<HostLev> H;
  H.push_back(std::move(L0));
  while (H.back().na > 1 || (coupled && H.back().nm > 1)) H.push_back(coarsen(H.back(), coupled));
  const HostLev &Lc = H.back();
  P.pin_s = -1;
  for (int s2 = ns - 1; s2 >= 0 && P.pin_s < 0; s2--)
    if (Lc.act[Lc.at(0, 0, s2)]) P.pin_s = s2;

  // eigenvalues of the FFT axis (fft.f90:151-189, divided by dl^2 with
  // dl = dxc_g(lo_g-1), main.f90:717-728)
  std::vector<double> lamh;
  if (!coupled) {
    const Block &B0 = cs.blocks[0];
    int kf, kb;
    double nr[2];
    if (!find_fft(cbc(B0, fa, 0), cbc(B0, fa, 1), kf, kb, nr, fa_centred)) {
      P.why = std::string("unsupported BC pair along the FFT axis") + (fa_centred ? "" : " (staggered)");
      return;
    }
    eigenvalues(nm + qfft, cbc(B0, fa, 0), cbc(B0, fa, 1), lamh, fa_centred);
    lamh.resize(nm);
    const double dl = B0.df[fa][1];
    for (auto &l : lamh) l /= dl * dl;
    P.norm = 1.0 / (nr[0] * (nm + nr[1]));
    P.kind_fwd = kf;
    P.kind_bwd = kb;
  }
  P.ng = coupled ? 1 : nm;
  P.gsize = coupled ? (long)nm * na * ns : (long)na * ns;
  std::vector<int> sing(P.ng, 0);
  std::vector<double> nact(P.ng, nact_plane);  // active cells per group
  P.nact_plane = std::max(nact_plane, 1.0);
  P.lam_ref = 1.0;
  for (double l : lamh) P.lam_ref = std::max(P.lam_ref, std::fabs(l));
  P.nH = (P.G[P.ax_a] + 7) / 8 * 8;
  P.nW = (P.G[P.ax_s] + 7) / 8 * 8;
  P.any_singular = false;
  for (int g = 0; g < P.ng; g++) {
    // the Helmholtz operators are shifted by -1/alpha > 0: never singular
    sing[g] = comp < 0 && zero_rowsum && (coupled || lamh[g] == 0.0);
    P.any_singular = P.any_singular || sing[g];
  }
  if (P.any_singular && P.pin_s < 0) { P.why = "internal: empty coarsest line"; return; }
  P.tol = cs.hypre_tol;
  P.maxiter = cs.hypre_maxiter;
  P.colors = coupled ? 4 : 2;
  P.nlev = (int)H.size();
  if (device < 0) { P.ok = true; return; }

  // device arrays: the FP64 hierarchy and its FP32 twin (mixed precision)
  P.h64 = MGHier<double>();
  P.h32 = MGHier<float>();
  // 32 lanes per line (two lines per wavefront) up to 544 rows, else 64
  const int lpl = ns <= 32 * 17 ? 32 : 64;
  const int nch = ns >= 2 ? std::min(lpl, ns / 2) : 1;
  auto f32 = [](const std::vector<double> &v) { return std::vector<float>(v.begin(), v.end()); };
  if (!coupled) {
    P.h64.lam = dev_upload(P, lamh);
    P.h32.lam = dev_upload(P, f32(lamh));
    if (comp >= 0) {  // the eigenvalues without the shift, and the shifted ones the solves read
      P.lam_base = P.h64.lam;
      P.lam64 = dev_upload(P, lamh);
      P.lam32 = dev_upload(P, f32(lamh));
      P.h64.lam = P.lam64;
      P.h32.lam = P.lam32;
    }
  }
  for (size_t l = 0; l < H.size(); l++) {
    const HostLev &h = H[l];
    MGLev L;
    L.nm = h.nm; L.na = h.na; L.ns = h.ns;
    L.coupled = coupled;
    L.per_a = h.per_a; L.per_m = h.per_m;
    L.qm = coupled ? (long)h.na * h.ns : 0;
    L.dg = dev_upload(P, h.dg);
    L.w = dev_upload(P, h.w);
    L.cs = dev_upload(P, h.cs);
    L.ca = dev_upload(P, h.ca);
    L.cm = coupled ? dev_upload(P, h.cm) : nullptr;
    L.act = dev_upload(P, h.act);
    L.nch = nch;
    L.base = ns >= 2 ? ns / nch : 1;
    L.extra = ns >= 2 ? ns % nch : 0;
    L.lpl = lpl;
    L.lvar = -1;
    L.auni = 0;
    if (!coupled) {
      bool au = true;
      const size_t nsz = (size_t)h.ns;
      for (size_t q = nsz; au && q < (size_t)h.na * nsz; q++) {
        const size_t q0 = q % nsz;
        au = h.dg[q] == h.dg[q0] && h.w[q] == h.w[q0] && h.cs[q] == h.cs[q0] && h.ca[q] == h.ca[q0] &&
             h.act[q] == h.act[q0];
      }
      L.auni = au ? 1 : 0;
    }
    L.mlist = nullptr;
    L.mcount = nullptr;
    P.h64.lev.push_back(L);
    MGLevT<float> Lf;
    Lf.nm = L.nm; Lf.na = L.na; Lf.ns = L.ns;
    Lf.coupled = L.coupled; Lf.per_a = L.per_a; Lf.per_m = L.per_m; Lf.qm = L.qm;
    Lf.dg = dev_upload(P, f32(h.dg));
    Lf.w = dev_upload(P, f32(h.w));
    Lf.cs = dev_upload(P, f32(h.cs));
    Lf.ca = dev_upload(P, f32(h.ca));
    Lf.cm = coupled ? dev_upload(P, f32(h.cm)) : nullptr;
    Lf.act = L.act;
    Lf.nch = L.nch; Lf.base = L.base; Lf.extra = L.extra; Lf.lpl = L.lpl; Lf.lvar = -1; Lf.auni = L.auni;
    Lf.mlist = nullptr; Lf.mcount = nullptr;
    P.h32.lev.push_back(Lf);
    const size_t n = (size_t)h.nm * h.na * h.ns;
    if (l == 0) {
      P.x = dev_zeros<double>(P, n);
      P.r = dev_zeros<double>(P, n);
      P.z = dev_zeros<double>(P, n);
      P.p = dev_zeros<double>(P, n);
      P.q = dev_zeros<double>(P, n);
      P.pd[0] = P.p;
      P.pd[1] = P.q;
      P.pf[0] = dev_zeros<float>(P, n);
      P.pf[1] = dev_zeros<float>(P, n);
      P.zf = dev_zeros<float>(P, n);
      P.rf = dev_zeros<float>(P, n);
      P.h64.xl.push_back(P.z);
      P.h64.bl.push_back(P.r);
      P.h32.xl.push_back(P.zf);
      P.h32.bl.push_back(P.rf);
    } else {
      P.h64.xl.push_back(dev_zeros<double>(P, n));
      P.h64.bl.push_back(dev_zeros<double>(P, n));
      P.h32.xl.push_back(dev_zeros<float>(P, n));
      P.h32.bl.push_back(dev_zeros<float>(P, n));
    }
  }
  // fused tail: the levels with at most 32 columns (FFT path, modes independent)
  int tail0 = -1;
  const int nl = (int)H.size();
  if (!coupled) {
    for (int l = 0; l < nl; l++)
      if (H[l].na <= 32) { tail0 = l; break; }
    if (tail0 >= 0 && nl - tail0 > MG_TAIL_MAX) tail0 = nl - MG_TAIL_MAX;
    if (tail0 >= 0 && nl - tail0 < 2) tail0 = -1;  // nothing to fuse
  }
  auto set_tail = [&](auto &Hh) {
    Hh.tail0 = tail0;
    if (tail0 < 0) return;
    Hh.tail.nl = nl - tail0;
    int off = 0;
    for (int q = 0; q < Hh.tail.nl; q++) {
      Hh.tail.lev[q] = Hh.lev[tail0 + q];
      Hh.tail.x[q] = Hh.xl[tail0 + q];
      Hh.tail.b[q] = Hh.bl[tail0 + q];
      const int n = (Hh.lev[tail0 + q].na * Hh.lev[tail0 + q].ns + 3) / 4 * 4;  // 16-byte aligned blocks
      Hh.tail.ox[q] = off;
      Hh.tail.ob[q] = off + n;
      off += 2 * n;
    }
    Hh.tail.lds_elems = off;
  };
  set_tail(P.h64);
  set_tail(P.h32);
  P.nchunk = (int)std::max<long>(1, std::min<long>((P.gsize + 2047) / 2048, std::max(1, 8192 / P.ng)));
  P.part = dev_zeros<double>(P, (size_t)2 * P.ng * P.nchunk);
  P.bb = dev_zeros<double>(P, 2 * P.ng);
  P.rr = dev_zeros<double>(P, 2 * P.ng);
  P.pq = dev_zeros<double>(P, 2 * P.ng);
  P.rz[0] = dev_zeros<double>(P, 2 * P.ng);
  P.rz[1] = dev_zeros<double>(P, 2 * P.ng);
  P.sums = dev_zeros<double>(P, 2 * P.ng);
  P.nact = dev_upload(P, nact);
  P.done = dev_zeros<int>(P, P.ng);
  P.singular = dev_upload(P, sing);
  {
    std::vector<int> sl;
    for (int g = 0; g < P.ng; g++)
      if (sing[g]) sl.push_back(g);
    P.n_sing = (int)sl.size();
    P.sing_list = dev_upload(P, sl);
  }
  // every group iterating until the first mg_flags of a solve says otherwise
  P.left = dev_upload(P, std::vector<int>{P.ng, 0});
  if (!coupled) {
    std::vector<int> ident(P.ng);
    for (int g = 0; g < P.ng; g++) ident[g] = g;
    P.mlist = dev_upload(P, ident);
    for (auto &L : P.h64.lev) { L.mlist = P.mlist; L.mcount = P.left; }
    for (auto &L : P.h32.lev) { L.mlist = P.mlist; L.mcount = P.left; }
    set_tail(P.h64);
    set_tail(P.h32);
  }
  SNAC_HIP(hipHostMalloc((void **)&P.left_host, 2 * sizeof(int), hipHostMallocDefault));
  P.stall = dev_zeros<int>(P, 1);
  P.rr_prev = dev_zeros<double>(P, 2 * P.ng);
  SNAC_HIP(hipHostMalloc((void **)&P.stall_host, sizeof(int), hipHostMallocDefault));
  P.ok = true;
}

namespace {

template <typename T>
void vcycle(MGPlan &P, MGHier<T> &H, hipStream_t st, size_t l) {
  const MGLevT<T> &L = H.lev[l];
  T *x = H.xl[l];
  const T *b = H.bl[l];
  if ((int)l == H.tail0) {  // the coarse end in one launch, one workgroup per mode
    if (P.mark_tail) P.mark_tail(true);
    mg_tail(st, H.tail, P.ng, H.lam, P.pin_s, P.singular, P.done);
    if (P.mark_tail) P.mark_tail(false);
    return;
  }
  if (l + 1 == H.lev.size()) {  // one line per group: direct solve
    mg_line(st, L, H.lam, x, b, 0, 0, 1, P.pin_s, P.singular, P.done);
    return;
  }
  int col[4][2], nc = 0;
  const bool two_a = L.na > 1, two_m = L.coupled && L.nm > 1;
  for (int cm = 0; cm <= (two_m ? 1 : 0); cm++)
    for (int ca = 0; ca <= (two_a ? 1 : 0); ca++) { col[nc][0] = ca; col[nc][1] = cm; nc++; }
  const bool zero_first = nc > 2;
  if (zero_first) mg_zero(st, (long)L.nm * L.na * L.ns, x);
  for (int k = 0; k < nc; k++) {
    const int znb = (!zero_first && k == 0) ? 1 : 0;
    const bool mk = l == 0 && znb == 0 && P.mark_line0;
    if (mk) P.mark_line0(true);
    mg_line(st, L, H.lam, x, b, col[k][0], col[k][1], znb, -1, nullptr, P.done);
    if (mk) P.mark_line0(false);
  }
  // smooth_only: the finest level's symmetric zebra sweeps are the whole
  // preconditioner (no coarse correction)
  const bool coarse = !(l == 0 && P.smooth_only);
  if (coarse) {
    mg_restrict_residual(st, L, H.lev[l + 1], H.lam, x, b, H.bl[l + 1], P.done);
    vcycle(P, H, st, l + 1);
  }
  // two colours (FFT path): the first post-sweep relaxes colour 1 reading
  // the colour-0 lines with the prolongation added on the fly, and colour
  // 0 is then relaxed from scratch, so the prolongation is never stored.
  // The coupled colours of the 3D path read not-yet-relaxed lines twice:
  // prolong first.
  const bool fuse = coarse && nc == 2 && !L.coupled;
  if (coarse && !fuse) mg_prolong(st, L, H.lev[l + 1], (const T *)H.xl[l + 1], x, P.done);
  for (int k = nc - 1; k >= 0; k--) {
    const bool mk = l == 0 && P.mark_line0;
    if (mk) P.mark_line0(true);
    const bool pro = fuse && k == nc - 1;
    mg_line(st, L, H.lam, x, b, col[k][0], col[k][1], 0, -1, nullptr, P.done, pro ? (const T *)H.xl[l + 1] : nullptr,
            pro ? H.lev[l + 1].na : 0);
    if (mk) P.mark_line0(false);
  }
}

template <typename T>
void project(MGPlan &P, hipStream_t st, T *v) {
  if (!P.any_singular) return;
  mg_dot<T, T>(st, P.ng, P.gsize, nullptr, nullptr, v, nullptr, P.part, P.nchunk, P.sums, P.done, P.singular);
  mg_sub_mean(st, P.h64.lev[0], v, P.sums, P.sing_list, P.n_sing, P.gsize, P.nact, P.done);
}

// z = V-cycle(r) (or the network's answer, net) with its null-space
// component removed; rz = r.z.  rr: the per-group r.r of the current
// residual (the network's input scale)
void precondition(MGPlan &P, hipStream_t st, double *rz_new, bool net, const double *rr) {
  if (P.prec >= 1) {
    if (net) {
      const MGLev &L0 = P.h64.lev[0];
      launch_net_in(st, P.ng, L0.na, L0.ns, P.nH, P.nW, P.rf, rr, P.nact_plane, P.h64.lam, P.lam_ref,
                    P.net->input());
      P.net->forward(st, P.ng);
      launch_net_out(st, P.ng, L0.na, L0.ns, P.nH, P.nW, P.net->output(), rr, P.nact_plane, L0.act, P.zf);
    } else {
      vcycle(P, P.h32, st, 0);
    }
    project(P, st, P.zf);
    mg_dot<float, float>(st, P.ng, P.gsize, P.r, P.zf, nullptr, nullptr, P.part, P.nchunk, rz_new, P.done, nullptr);
  } else {
    vcycle(P, P.h64, st, 0);
    project(P, st, P.z);
    mg_dot<double, double>(st, P.ng, P.gsize, P.r, P.z, nullptr, nullptr, P.part, P.nchunk, rz_new, P.done, nullptr);
  }
}

}  // namespace

namespace {

// one CG iteration after the first: precondition the current residual, then
// the operator application and the update with their dot products and the
// convergence flags
void mg_iteration(MGPlan &P, hipStream_t st, MGRun &R) {
  const bool mixed = P.prec >= 1;
  const double *rzo = nullptr;
  if (R.it > 0) {
    precondition(P, st, P.rz[1 - R.cur], R.net, P.rr);
    rzo = P.rz[R.cur];
    R.cur = 1 - R.cur;
  }
  // p = z + beta p (first iteration: p = z) with p.(M p); beta of the
  // flexible form (FP32 preconditioners) takes r_new . z_old off rz_new.
  // Then x += alpha p, r -= alpha M p with r.r and, for the flexible form,
  // r_new . z_old in the second slot.  M p is applied twice rather than
  // stored: the second pass recomputes it from p.
  // The V-cycle's output is well scaled and its search directions are kept
  // in FP32; the network's need not be (an untrained one's directions grow
  // past the FP32 range), so the learned mode keeps FP64 directions.
  const MGLev &L0 = P.h64.lev[0];
  if (mixed && P.prec == 1) {
    mg_dir_apply_dot<float, float>(st, L0, P.h64.lam, P.ng, P.gsize, P.rz[R.cur], rzo, rzo ? P.rr : nullptr, P.zf,
                                   P.pf[R.pc], P.pf[1 - R.pc], P.part, P.nchunk, P.pq, P.done);
    R.pc = 1 - R.pc;
    if (P.mark_update) P.mark_update(true);
    mg_update_apply_dot<float>(st, L0, P.h64.lam, P.ng, P.gsize, P.rz[R.cur], P.pq, P.x, P.pf[R.pc], P.r, P.rf, P.zf,
                               P.part, P.nchunk, P.rr, P.done, R.x_unset);
    if (P.mark_update) P.mark_update(false);
  } else if (mixed) {
    mg_dir_apply_dot<float, double>(st, L0, P.h64.lam, P.ng, P.gsize, P.rz[R.cur], rzo, rzo ? P.rr : nullptr, P.zf,
                                    P.pd[R.pc], P.pd[1 - R.pc], P.part, P.nchunk, P.pq, P.done);
    R.pc = 1 - R.pc;
    if (P.mark_update) P.mark_update(true);
    mg_update_apply_dot<double>(st, L0, P.h64.lam, P.ng, P.gsize, P.rz[R.cur], P.pq, P.x, P.pd[R.pc], P.r, P.rf,
                                P.zf, P.part, P.nchunk, P.rr, P.done, R.x_unset);
    if (P.mark_update) P.mark_update(false);
  } else {
    mg_dir_apply_dot<double, double>(st, L0, P.h64.lam, P.ng, P.gsize, P.rz[R.cur], rzo, nullptr, P.z, P.pd[R.pc],
                                     P.pd[1 - R.pc], P.part, P.nchunk, P.pq, P.done);
    R.pc = 1 - R.pc;
    if (P.mark_update) P.mark_update(true);
    mg_update_apply_dot<double>(st, L0, P.h64.lam, P.ng, P.gsize, P.rz[R.cur], P.pq, P.x, P.pd[R.pc], P.r, nullptr,
                                nullptr, P.part, P.nchunk, P.rr, P.done, R.x_unset);
    if (P.mark_update) P.mark_update(false);
  }
  R.x_unset = false;  // written by this iteration's update pass
  mg_flags(st, P.ng, P.rr, P.bb, P.tol * P.tol, P.done, P.left, P.mlist, R.it + 1);
  if (R.net && P.fallback) launch_net_stall(st, P.ng, P.rr, P.rr_prev, P.done, P.stall_ratio * P.stall_ratio, P.stall);
  R.it++;
}

// groups still above the tolerance (synchronises); left_host[1] is the
// iteration count at which none was left (0: not reached yet)
int groups_left(MGPlan &P, hipStream_t st) {
  SNAC_HIP(hipMemcpyAsync(P.left_host, P.left, 2 * sizeof(int), hipMemcpyDeviceToHost, st));
  SNAC_HIP(hipStreamSynchronize(st));
  return *P.left_host;
}

}  // namespace

MGRun mg_solve(MGPlan &P, hipStream_t st, int fixed) {
  const long n = (long)P.ng * P.gsize;
  const double tol2 = P.tol * P.tol;
  const bool mixed = P.prec >= 1;  // the flexible CG of the FP32 preconditioners
  const bool planar = P.fa >= 0;   // one 2D plane per group (the network's input)
  MGRun R;
  R.net = P.prec == 2 && P.net && planar && fixed == 0;
  R.record = P.collecting && planar && fixed == 0 && P.pair_n + P.ng <= P.pair_cap;
  const MGLev &L0 = P.h64.lev[0];
  const size_t pl = (size_t)P.nH * P.nW;
  // the warm start keeps x (the previous solve's solution) as x0; the
  // network's modes start from zero or from its own guess
  const bool warm = P.warm && !R.net && !R.record;
  // starting from zero: the first update pass writes x = alpha p without
  // reading x, and zeroes the groups already done (the network's initial
  // guess adds to a zeroed x instead)
  R.x_unset = !warm && !(R.net && P.guess) && P.maxiter > 0;
  if (!warm && !R.x_unset) mg_zero(st, n, P.x);
  mg_izero(st, P.ng, P.done);
  mg_izero(st, 1, P.left + 1);
  project(P, st, P.r);
  // b.b, and the FP32 copy of r in the same pass where it is needed as is
  const bool cvt = mixed && !warm;
  if (cvt) mg_dot_self_cvt(st, P.ng, P.gsize, P.r, P.rf, P.part, P.nchunk, P.bb);
  else mg_dot<double, double>(st, P.ng, P.gsize, P.r, P.r, nullptr, nullptr, P.part, P.nchunk, P.bb, P.done, nullptr);
  mg_flags(st, P.ng, P.bb, P.bb, tol2, P.done, P.left, P.mlist, 0);
  if (R.record)  // the input half of this solve's training pairs
    launch_net_pair_in(st, P.ng, L0.na, L0.ns, P.nH, P.nW, P.r, P.bb, P.nact_plane, P.h64.lam, P.lam_ref,
                       P.pair_x + (size_t)P.pair_n * pl * 2);
  if (warm) {  // r = b - M x0; groups that x0 already solves stop here
    mg_apply_dot(st, L0, P.h64.lam, P.x, P.q, P.ng, P.gsize, P.part, P.nchunk, P.pq, P.done);
    mg_warm_residual(st, P.ng, P.gsize, P.bb, P.x, P.r, P.q, mixed ? P.rf : nullptr, P.part, P.nchunk, P.rr);
    mg_flags(st, P.ng, P.rr, P.bb, tol2, P.done, P.left, P.mlist, 0);
  }
  const double *rr0 = P.bb;  // r.r of the residual the first preconditioning sees
  if (R.net && P.guess) {
    // initial guess x0 = the network's answer for b, then r = b - M x0 (the
    // CG update with p = x0 and the step length rz / pq forced to 1)
    launch_net_in(st, P.ng, L0.na, L0.ns, P.nH, P.nW, P.rf, P.bb, P.nact_plane, P.h64.lam, P.lam_ref, P.net->input());
    P.net->forward(st, P.ng);
    launch_net_out(st, P.ng, L0.na, L0.ns, P.nH, P.nW, P.net->output(), P.bb, P.nact_plane, L0.act, P.zf);
    project(P, st, P.zf);
    mg_cg_dir<float>(st, P.ng, P.gsize, P.rz[0], nullptr, nullptr, P.zf, P.p, P.done);
    mg_apply_dot(st, L0, P.h64.lam, P.p, P.q, P.ng, P.gsize, P.part, P.nchunk, P.pq, P.done);
    mg_cg_update_dot(st, P.ng, P.gsize, P.pq, P.pq, P.x, P.p, P.r, P.q, P.rf, nullptr, P.part, P.nchunk, P.rr,
                     P.done);
    mg_flags(st, P.ng, P.rr, P.bb, tol2, P.done, P.left, P.mlist, 0);
    rr0 = P.rr;
  }
  if (R.net) {
    mg_izero(st, 1, P.stall);
    SNAC_HIP(hipMemcpyAsync(P.rr_prev, rr0, 2 * P.ng * sizeof(double), hipMemcpyDeviceToDevice, st));
  }
  precondition(P, st, P.rz[R.cur], R.net, rr0);
  if (fixed > 0) {  // captured: counted when the host resumes it (mg_continue)
    const int nit = std::min(fixed, P.maxiter);
    while (R.it < nit) mg_iteration(P, st, R);
    return R;
  }
  P.n_solves++;
  while (R.it < P.maxiter) {
    mg_iteration(P, st, R);
    if (R.net && P.fallback) SNAC_HIP(hipMemcpyAsync(P.stall_host, P.stall, sizeof(int), hipMemcpyDeviceToHost, st));
    if (groups_left(P, st) == 0) break;
    if (R.net && P.fallback && *P.stall_host) {  // the network stopped helping: V-cycles from here on
      R.net = false;
      P.fallbacks++;
    }
  }
  if (*P.left_host > 0) P.n_unconverged++;
  if (R.record) {  // the target half: the solution, scaled like the input
    launch_net_pair_out(st, P.ng, L0.na, L0.ns, P.nH, P.nW, P.x, P.bb, P.nact_plane, P.pair_t + (size_t)P.pair_n * pl);
    P.pair_n += P.ng;
  }
  P.last_iters = R.it;
  P.max_seen = std::max(P.max_seen, P.last_iters);
  return R;
}

void mg_continue(MGPlan &P, hipStream_t st, MGRun &R) {
  P.n_solves++;
  int left = groups_left(P, st);
  if (left > 0 && R.it < P.maxiter) P.n_continued++;
  while (left > 0 && R.it < P.maxiter) {
    mg_iteration(P, st, R);
    left = groups_left(P, st);
  }
  if (left > 0) P.n_unconverged++;
  // the iterations the solve needed: the captured ones may have run past the
  // point where every group met the tolerance (those groups stay frozen)
  P.last_iters = P.left_host[1] > 0 ? P.left_host[1] - 1 : R.it;
  P.max_seen = std::max(P.max_seen, P.last_iters);
}

void mg_stats(MGPlan &P, hipStream_t st, int &iters, double &max_relres) {
  std::vector<double> rr(2 * P.ng), bb(2 * P.ng);
  SNAC_HIP(hipMemcpyAsync(rr.data(), P.rr, rr.size() * sizeof(double), hipMemcpyDeviceToHost, st));
  SNAC_HIP(hipMemcpyAsync(bb.data(), P.bb, bb.size() * sizeof(double), hipMemcpyDeviceToHost, st));
  SNAC_HIP(hipStreamSynchronize(st));
  iters = P.last_iters;
  max_relres = 0.0;
  for (int g = 0; g < P.ng; g++) {
    if (std::isnan(rr[2 * g]) || std::isnan(bb[2 * g])) {  // a broken solve never reads as converged
      max_relres = std::nan("");
      break;
    }
    if (bb[2 * g] > 0.0) max_relres = std::max(max_relres, std::sqrt(rr[2 * g] / bb[2 * g]));
  }
}

}  // namespace snac

namespace snac {
namespace {

// The line kernels of a hierarchy's levels, per hierarchy shape, kept for
// the life of the process: a later plan of the same shapes reuses the
// choices without timing, so that same-shape contexts run the same kernels
// and agree bit for bit (variants 3 and 4 round differently from 0..2).
std::mutex g_line_mu;
std::map<std::string, std::vector<int>> g_line_choice;

template <typename T>
std::string hier_key(const MGHier<T> &H, int dev) {
  std::ostringstream k;
  k << "dev " << dev << " T" << sizeof(T) << " tail " << H.tail0;
  for (const auto &L : H.lev)
    k << " | nm " << L.nm << " na " << L.na << " ns " << L.ns << " qm " << L.qm << " c " << L.coupled << " p "
      << L.per_a << " lpl " << L.lpl;
  return k.str();
}

// Chooses each launched level's line kernel by timing whole V-cycles (the
// sweeps and restrictions of every level in their real order and cache
// state; a level's kernel timed alone, back to back on cache-resident
// vectors, ranks the variants differently): first the fastest variant for
// all levels at once, then each level's own variant in turn with the others
// fixed, a change kept only when it saves more than 2 %.
template <typename T>
void tune_hier(MGPlan &P, MGHier<T> &H, hipStream_t st) {
  const int nl = (int)H.lev.size();
  const int last = H.tail0 >= 0 ? H.tail0 : nl - 1;  // launched levels that relax lines and restrict
  int dev = 0;
  SNAC_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_line_mu);
  const std::string key = hier_key(H, dev);
  auto it = g_line_choice.find(key);
  if (it != g_line_choice.end()) {
    for (int l = 0; l < nl; l++) H.lev[l].lvar = it->second[l];
    return;
  }
  std::vector<int> lv(nl, -1);
  std::vector<int> tuned;  // the launched levels with a choice to make
  for (int l = 0; l < last; l++)
    if (H.lev[l].na >= 2) tuned.push_back(l);
  if (!tuned.empty()) {
    hipEvent_t e0, e1;
    SNAC_HIP(hipEventCreate(&e0));
    SNAC_HIP(hipEventCreate(&e1));
    auto apply = [&](const std::vector<int> &v) {
      for (int l = 0; l < nl; l++) H.lev[l].lvar = v[l];
    };
    // a right-hand side with values (zeros would leave every line trivial)
    mg_pattern(st, (long)H.lev[0].nm * H.lev[0].na * H.lev[0].ns, H.bl[0]);
    auto time_cycles = [&](const std::vector<int> &v) {
      apply(v);
      vcycle(P, H, st, 0);  // warm-up
      SNAC_HIP(hipEventRecord(e0, st));
      for (int r = 0; r < 3; r++) vcycle(P, H, st, 0);
      SNAC_HIP(hipEventRecord(e1, st));
      SNAC_HIP(hipEventSynchronize(e1));
      float ms = 0.f;
      SNAC_HIP(hipEventElapsedTime(&ms, e0, e1));
      return ms;
    };
    // the variant a level runs when v is asked for (the next lower one it allows)
    auto runs_as = [&](int l, int v) {
      while (v > 0 && !mg_line_variant_runs(H.lev[l], v)) v--;
      return v;
    };
    // untimed V-cycles first: the clocks of an idle GPU ramp up under load,
    // and latency-bound and throughput-bound kernels gain differently
    apply(lv);
    for (int r = 0; r < 30; r++) vcycle(P, H, st, 0);
    // each candidate timed in two rounds, the better of the two kept
    float best = 1e30f;
    std::vector<int> bestv = lv;
    std::vector<std::vector<int>> cands;
    for (int v = 0; v <= 4; v++) {
      std::vector<int> cand = lv;
      bool any = false;
      for (int l : tuned) {
        cand[l] = runs_as(l, v);
        any = any || cand[l] == v;
      }
      if (any) cands.push_back(cand);  // else every level would fall back to a variant already timed
    }
    std::vector<float> ms(cands.size(), 1e30f);
    for (int round = 0; round < 2; round++)
      for (size_t c = 0; c < cands.size(); c++) ms[c] = std::min(ms[c], time_cycles(cands[c]));
    for (size_t c = 0; c < cands.size(); c++)
      if (ms[c] < best) { best = ms[c]; bestv = cands[c]; }
    for (int l : tuned) {
      for (int v = 0; v <= 4; v++) {
        if (v == bestv[l] || runs_as(l, v) != v) continue;
        std::vector<int> cand = bestv;
        cand[l] = v;
        const float t = std::min(time_cycles(cand), time_cycles(cand));
        if (t < 0.98f * best) { best = t; bestv = cand; }
      }
    }
    lv = bestv;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
  for (int l = 0; l < nl; l++) H.lev[l].lvar = lv[l];
  g_line_choice.emplace(key, lv);
  // the trials ran V-cycles on a patterned right-hand side: zero the vectors again
  for (int l = 0; l < nl; l++) {
    const long n = (long)H.lev[l].nm * H.lev[l].na * H.lev[l].ns;
    mg_zero(st, n, H.xl[l]);
    mg_zero(st, n, H.bl[l]);
  }
  SNAC_HIP(hipStreamSynchronize(st));
}

}  // namespace

void mg_tune_lines(MGPlan &P, hipStream_t st) {
  if (!P.ok || !P.x) return;
  tune_hier(P, P.h64, st);
  tune_hier(P, P.h32, st);
}

}  // namespace snac
