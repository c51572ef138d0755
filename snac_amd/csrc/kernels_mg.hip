// Kernels of the batched MG-preconditioned CG (mg.hpp): zebra line
// relaxation with one wavefront per line (partitioned Thomas over the lanes'
// row chunks, parallel cyclic reduction of the 2-per-lane interface system
// with cross-lane shuffles), fused residual + restriction, prolongation,
// operator application, per-group dot products with wavefront reductions and
// the CG vector updates.
#include "mg.hpp"
#include <algorithm>
#include <type_traits>

namespace snac {
namespace {

template <typename T>
__device__ __forceinline__ long vidx(const MGLevT<T> &L, int m, int a, int s) { return ((long)m * L.na + a) * L.ns + s; }
template <typename T>
__device__ __forceinline__ long cidx(const MGLevT<T> &L, int m, int a, int s) { return m * L.qm + (long)a * L.ns + s; }

// sum of the couplings to the neighbouring lines (a and m directions) times x
template <typename T>
__device__ __forceinline__ T nb_sum(const MGLevT<T> &L, const T *__restrict__ x, int m, int a, long e, long q) {
  T r = 0;
  const long sa = L.ns;
  if (a < L.na - 1) r += L.ca[q] * x[e + sa];
  else if (L.per_a) r += L.ca[q] * x[e - (long)(L.na - 1) * sa];
  if (a > 0) r += L.ca[q - sa] * x[e - sa];
  else if (L.per_a) r += L.ca[q + (long)(L.na - 1) * sa] * x[e + (long)(L.na - 1) * sa];
  if (L.coupled) {
    const long sm = (long)L.na * L.ns;
    if (m < L.nm - 1) r += L.cm[q] * x[e + sm];
    else if (L.per_m) r += L.cm[q] * x[e - (long)(L.nm - 1) * sm];
    if (m > 0) r += L.cm[q - L.qm] * x[e - sm];
    else if (L.per_m) r += L.cm[q + (long)(L.nm - 1) * L.qm] * x[e + (long)(L.nm - 1) * sm];
  }
  return r;
}

// nb_sum of an FFT-path level (no m couplings) whose neighbour lines still
// lack the prolongation: each neighbour reads x + xc of its coarse cell
// (active cells only), the value the separate prolongation pass would have
// stored there (xc: the next coarser level's x, cna columns)
template <typename T>
__device__ __forceinline__ T nb_sum_pro(const MGLevT<T> &L, const T *__restrict__ x, int m, int a, int s, long e,
                                        long q, const T *__restrict__ xc, int cna) {
  const long sa = L.ns;
  const int ra = cna != L.na;
  auto v = [&](int an, long en, long qn) -> T {
    T xv = x[en];
    if (L.act[qn]) xv += xc[((long)m * cna + (ra ? an >> 1 : an)) * sa + s];
    return xv;
  };
  const long w = (long)(L.na - 1) * sa;
  T r = 0;
  if (a < L.na - 1) r += L.ca[q] * v(a + 1, e + sa, q + sa);
  else if (L.per_a) r += L.ca[q] * v(0, e - w, q - w);
  if (a > 0) r += L.ca[q - sa] * v(a - 1, e - sa, q - sa);
  else if (L.per_a) r += L.ca[q + w] * v(L.na - 1, e + w, q + w);
  return r;
}

template <typename T>
__device__ __forceinline__ T mx_at(const MGLevT<T> &L, T lm, const T *__restrict__ x, int m, int a, int s) {
  const long e = vidx(L, m, a, s), q = cidx(L, m, a, s);
  T y = (L.dg[q] - lm * L.w[q]) * x[e];
  if (s > 0) y -= L.cs[q - 1] * x[e - 1];
  if (s < L.ns - 1) y -= L.cs[q] * x[e + 1];
  return y - nb_sum(L, x, m, a, e, q);
}

template <typename T>
__device__ __forceinline__ int group_of(const MGLevT<T> &L, int m) { return L.coupled ? 0 : m; }

// (m, a, s) of the element at flat index idx of level L
template <typename T>
__device__ __forceinline__ void pos_of(const MGLevT<T> &L, long idx, int &m, int &a, int &s) {
  s = (int)(idx % L.ns);
  const long t = idx / L.ns;
  a = (int)(t % L.na);
  m = (int)(t / L.na);
}
// advances (m, a, s) by da * ns + ds elements (0 <= ds < ns): a grid-stride
// loop steps its position without dividing the flat index every element
template <typename T>
__device__ __forceinline__ void pos_step(const MGLevT<T> &L, int da, int ds, int &m, int &a, int &s) {
  s += ds;
  a += da;
  if (s >= L.ns) { s -= L.ns; a++; }
  while (a >= L.na) { a -= L.na; m++; }
}

template <typename T>
__device__ __forceinline__ T shfl(T v, int src) { return __shfl(v, src, 64); }

// LDS position of row s of a staged line: one pad word after every 32 rows,
// so that the lanes of a segment, which read rows base apart (base a
// multiple of 8 on power-of-two lines), fall on different banks
__host__ __device__ __forceinline__ int lpad(int s) { return s + (s >> 5); }
// LDS words per staged line of ns rows
__host__ __device__ __forceinline__ int lpad_len(int ns) { return lpad(ns - 1) + 1; }

// One reduced row  al*u[i-d] + be*u[i] + ga*u[i+d] = de  of the interface
// system; rows outside the line are the identity.
template <typename T>
struct Row {
  T al, be, ga, de;
};

template <typename T>
__device__ __forceinline__ Row<T> fetch(const Row<T> &r, int src, bool valid) {
  Row<T> o;
  o.al = shfl(r.al, src);
  o.be = shfl(r.be, src);
  o.ga = shfl(r.ga, src);
  o.de = shfl(r.de, src);
  if (!valid) { o.al = 0; o.be = 1; o.ga = 0; o.de = 0; }
  return o;
}

// Reciprocals for the relaxation's eliminations without the IEEE division
// sequence: the hardware estimate refined by Newton steps (two for FP64,
// one for FP32), within a few ulp of the correctly rounded quotient.  The
// relaxation only preconditions a CG that the FP64 residual judges.
__device__ __forceinline__ double frcp(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = __builtin_fma(-x, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-x, r, 1.0);
  return __builtin_fma(r, e, r);
}
__device__ __forceinline__ float frcp(float x) {
  float r = __builtin_amdgcn_rcpf(x);
  float e = __builtin_fmaf(-x, r, 1.0f);
  return __builtin_fmaf(r, e, r);
}
// a / b, or a times the refined reciprocal of b (FR)
template <bool FR, typename T>
__device__ __forceinline__ T dv(T a, T b) {
  if constexpr (FR) return a * frcp(b);
  else return a / b;
}

// PCR elimination of row r against its neighbours lo (index i-d) and hi (i+d)
template <bool FR, typename T>
__device__ __forceinline__ Row<T> pcr(const Row<T> &r, const Row<T> &lo, const Row<T> &hi) {
  const T k1 = dv<FR>(r.al, lo.be), k2 = dv<FR>(r.ga, hi.be);
  Row<T> o;
  o.al = -k1 * lo.al;
  o.be = r.be - k1 * lo.ga - k2 * hi.al;
  o.ga = -k2 * hi.ga;
  o.de = r.de - k1 * lo.de - k2 * hi.de;
  return o;
}

// Line relaxation: one wavefront per line (m, a), 4 lines per workgroup.
// Lane c owns rows [p0, p0+cnt) of the line.  Each lane eliminates its
// chunk's interior (Thomas downwards, then upwards) leaving two interface
// rows per lane that form a tridiagonal system in (x_first, x_last) of all
// chunks; PCR over the lanes solves it in log2(2*nch) steps; the interior
// rows follow from the two interface values.
// LPL lanes per line (32 or 64): a wavefront relaxes 64 / LPL lines at once,
// each in its own lane segment (the shuffles never cross segments); act =
// false leaves the segment idle.

// the elimination of one line whose rows [p0, p0 + cnt) belong to the
// calling lane: row(j, l, d, u, r) gives row p0 + j (j < cnt; each row is
// asked for once, and before any put), put(j, v) stores its solution
// (cnt = 0: an idle lane, which still takes part in the shuffles)
// FULL: every lane of the segment holds exactly K rows (ns = LPL * K), so the
// row guards and the interface system's size are compile-time constants
template <int K, int LPL, typename T, bool FR, bool FULL = false, typename ROW, typename PUT>
__device__ __forceinline__ void line_elim(int ns, int nch, int cnt, bool act, ROW &&row, PUT &&put) {
  if constexpr (FULL) {
    ns = LPL * K;
    nch = LPL;
    cnt = K;
  }
  const int wl = threadIdx.x & 63, lane = wl & (LPL - 1), seg0 = wl - lane;
  if (ns == 1) {  // a one-row line: no interface system
    if (act && lane == 0) {
      T l, d, u, r;
      row(0, l, d, u, r);
      put(0, dv<FR>(r, d));
    }
    return;
  }
  // downward elimination: row j (1 <= j < cnt) becomes A x_p0 + x_j + C x_{j+1} = D
  T A[K], C[K], D[K];
  T aq = 0, cq = 0, dq = 0;
#pragma unroll
  for (int j = 0; j < K; j++) { A[j] = 0; C[j] = 0; D[j] = 0; }
#pragma unroll
  for (int j = 1; j < K; j++) {
    if (j < cnt) {
      T lo, dg, up, rh;
      row(j, lo, dg, up, rh);
      if (j == 1) {
        const T ir = dv<FR>(T(1), dg);
        A[1] = lo * ir; C[1] = up * ir; D[1] = rh * ir;
      } else {
        const T ir = dv<FR>(T(1), dg - lo * C[j - 1]);
        A[j] = -lo * A[j - 1] * ir; C[j] = up * ir; D[j] = (rh - lo * D[j - 1]) * ir;
      }
      if (j == cnt - 1) { aq = A[j]; cq = C[j]; dq = D[j]; }
    }
  }
  // upward elimination: x_j = D - A x_p0 - C x_last for 1 <= j <= cnt-2
  T dn = 0, an = 0, cn = 0;
#pragma unroll
  for (int j = K - 2; j >= 1; j--) {
    if (j <= cnt - 2) {
      if (j == cnt - 2) { dn = D[j]; an = A[j]; cn = C[j]; }
      else { dn = D[j] - C[j] * dn; an = A[j] - C[j] * an; cn = -C[j] * cn; }
      D[j] = dn; A[j] = an; C[j] = cn;
    }
  }
  Row<T> r0, r1;
  if (cnt >= 2) {
    T lo, dg, up, rh;
    row(0, lo, dg, up, rh);
    r0.al = lo;
    if (cnt == 2) { r0.be = dg; r0.ga = up; r0.de = rh; }
    else { r0.be = dg - up * an; r0.ga = -up * cn; r0.de = rh - up * dn; }
    r1.al = aq; r1.be = 1; r1.ga = cq; r1.de = dq;
  } else {
    r0 = Row<T>{0, 1, 0, 0};
    r1 = r0;
  }
  // interface system: u[2c] = x_first(c), u[2c+1] = x_last(c); PCR, d = 1
  {
    const Row<T> r1m = fetch(r1, seg0 + (lane > 0 ? lane - 1 : 0), lane > 0);
    const Row<T> r0p = fetch(r0, seg0 + (lane < LPL - 1 ? lane + 1 : LPL - 1), lane < LPL - 1);
    const Row<T> n0 = pcr<FR>(r0, r1m, r1);
    const Row<T> n1 = pcr<FR>(r1, r0, r0p);
    r0 = n0;
    r1 = n1;
  }
  const int nred = 2 * nch;
  for (int h = 1; 2 * h < nred; h <<= 1) {  // d = 2h: lanes -+h, same slot
    const bool vm = lane >= h, vp = lane + h < LPL;
    const int sm = seg0 + (vm ? lane - h : lane), sp = seg0 + (vp ? lane + h : lane);
    const Row<T> a0 = fetch(r0, sm, vm), b0 = fetch(r0, sp, vp);
    const Row<T> a1 = fetch(r1, sm, vm), b1 = fetch(r1, sp, vp);
    r0 = pcr<FR>(r0, a0, b0);
    r1 = pcr<FR>(r1, a1, b1);
  }
  if (cnt == 0) return;
  const T xp = dv<FR>(r0.de, r0.be), xq = dv<FR>(r1.de, r1.be);
  put(0, xp);
  put(cnt - 1, xq);
#pragma unroll
  for (int j = 1; j < K - 1; j++)
    if (j <= cnt - 2) put(j, D[j] - A[j] * xp - C[j] * xq);
}

// one row of line (m, a): diagonal, couplings below and above and the
// right-hand side with the neighbour lines' contribution (zero_nb: read as
// zero; xc: with the prolongation of xc added), or the identity row of the
// pinned row of a singular group
template <typename T>
__device__ __forceinline__ void line_row(const MGLevT<T> &L, T lmv, const T *__restrict__ x, const T *__restrict__ b,
                                         int m, int a, int s, long e, long q, int zero_nb, int pin_s, bool pin,
                                         const T *__restrict__ xc, int cna, T &l, T &d, T &u, T &r) {
  d = L.dg[q] - lmv * L.w[q];
  l = s > 0 ? -L.cs[q - 1] : T(0);
  u = s < L.ns - 1 ? -L.cs[q] : T(0);
  r = b[e];
  if (!zero_nb) r += xc ? nb_sum_pro(L, x, m, a, s, e, q, xc, cna) : nb_sum(L, x, m, a, e, q);
  if (pin && s == pin_s) { d = 1; l = 0; u = 0; r = 0; }
}

// relaxes line (m, a) with the calling lane segment, rows read from and x
// written to global memory directly (lane c's rows are contiguous)
// xc (FFT path only): the neighbour lines get the prolongation of the
// coarser level's xc (cna columns) added as they are read (nb_sum_pro)
// FULL: the line has exactly LPL * K rows (every lane K of them); an idle
// segment (act false) then leaves at once, its lanes' values unread by the
// other segments' shuffles
template <int K, int LPL, typename T, bool FR = false, bool FULL = false>
__device__ __forceinline__ void line_solve(const MGLevT<T> &L, T lmv, T *__restrict__ x, const T *__restrict__ b,
                                           int m, int a, int zero_nb, int pin_s, bool pin, bool act,
                                           const T *__restrict__ xc = nullptr, int cna = 0) {
  if (FULL && !act) return;
  const int wl = threadIdx.x & 63, lane = wl & (LPL - 1);
  const long e0 = vidx(L, m, a, 0), q0 = cidx(L, m, a, 0);
  const int cnt = FULL ? K : (act && lane < L.nch ? L.base + (lane < L.extra ? 1 : 0) : 0);
  const int p0 = FULL ? lane * K : lane * L.base + min(lane, L.extra);
  T lo[K], dg[K], up[K], rh[K];
#pragma unroll
  for (int j = 0; j < K; j++) {
    lo[j] = 0; dg[j] = 1; up[j] = 0; rh[j] = 0;
    if (j < cnt) {
      const int s = p0 + j;
      line_row(L, lmv, x, b, m, a, s, e0 + s, q0 + s, zero_nb, pin_s, pin, xc, cna, lo[j], dg[j], up[j], rh[j]);
    }
  }
  line_elim<K, LPL, T, FR, FULL>(
      L.ns, L.nch, cnt, act, [&](int j, T &l, T &d, T &u, T &r) { l = lo[j]; d = dg[j]; u = up[j]; r = rh[j]; },
      [&](int j, T v) { x[e0 + p0 + j] = v; });
}

template <int K, int LPL, typename T>
__global__ __launch_bounds__(256) void k_line(MGLevT<T> L, const T *__restrict__ lam, T *__restrict__ x,
                                              const T *__restrict__ b, int col_a, int col_m, int nla, int nlm,
                                              int zero_nb, int pin_s, const int *__restrict__ singular,
                                              const int *__restrict__ done, const T *__restrict__ xc, int cna) {
  const long line = ((long)blockIdx.x * 256 + threadIdx.x) / LPL;
  if (line >= (long)nla * nlm) return;  // a whole lane segment leaves together
  const int a = 2 * (int)(line % nla) + col_a;
  const int lm_ = (int)(line / nla);
  const int m = L.coupled ? 2 * lm_ + col_m : lm_;
  const int g = group_of(L, m);
  if (done && done[g]) return;
  const T lmv = lam ? lam[m] : T(0);
  const bool pin = pin_s >= 0 && singular && singular[g];
  line_solve<K, LPL, T>(L, lmv, x, b, m, a, zero_nb, pin_s, pin, true, xc, cna);
}

// The same relaxation with the right-hand sides staged through LDS: the
// workgroup's 256 / LPL lines first form b + (neighbour lines' couplings)
// row by row with consecutive threads on consecutive rows, so the streams
// of b, of the neighbour lines and of x are read and written whole rows at
// a time; each lane then takes its own contiguous rows from LDS, eliminates
// as line_solve does (the same arithmetic, so the same values), and x goes
// back through LDS to coalesced stores.  The coefficients (mode-independent
// on the FFT path, so cache-resident) are still read per lane.
template <int K, int LPL, typename T>
__global__ __launch_bounds__(256) void k_line_lds(MGLevT<T> L, const T *__restrict__ lam, T *__restrict__ x,
                                                  const T *__restrict__ b, int col_a, int col_m, int nla, int nlm,
                                                  int zero_nb, int pin_s, const int *__restrict__ singular,
                                                  const int *__restrict__ done, const T *__restrict__ xc, int cna) {
  constexpr int NLW = 256 / LPL;  // lines per workgroup
  extern __shared__ unsigned char smem_[];
  T *sh = reinterpret_cast<T *>(smem_);  // [NLW][nsp]: right-hand sides, then x (rows at lpad)
  const int ns = L.ns, nsp = lpad_len(ns);
  const long nl = (long)nla * nlm;
  const long line0 = (long)blockIdx.x * NLW;
  // phase 1: right-hand sides, one line after the other, rows across threads
  for (int li = 0; li < NLW; li++) {
    const long line = line0 + li;
    if (line >= nl) break;
    const int a = 2 * (int)(line % nla) + col_a;
    const int lm_ = (int)(line / nla);
    const int m = L.coupled ? 2 * lm_ + col_m : lm_;
    if (done && done[group_of(L, m)]) continue;
    const T lmv = lam ? lam[m] : T(0);
    const bool pin = pin_s >= 0 && singular && singular[group_of(L, m)];
    const long e0 = vidx(L, m, a, 0), q0 = cidx(L, m, a, 0);
    for (int s = threadIdx.x; s < ns; s += 256) {
      T l, d, u, r;
      line_row(L, lmv, x, b, m, a, s, e0 + s, q0 + s, zero_nb, pin_s, pin, xc, cna, l, d, u, r);
      sh[li * nsp + lpad(s)] = r;
    }
  }
  __syncthreads();
  // phase 2: each lane segment eliminates its line
  {
    const int li = threadIdx.x / LPL;
    const long line = line0 + li;
    bool act = line < nl;
    int m = 0, a = 0;
    if (act) {
      a = 2 * (int)(line % nla) + col_a;
      const int lm_ = (int)(line / nla);
      m = L.coupled ? 2 * lm_ + col_m : lm_;
      act = !(done && done[group_of(L, m)]);
    }
    // a whole segment shares act: idle segments skip the shuffles together
    if (act) {
      const T lmv = lam ? lam[m] : T(0);
      const bool pin = pin_s >= 0 && singular && singular[group_of(L, m)];
      const int lane = (threadIdx.x & 63) & (LPL - 1);
      const int cnt = lane < L.nch ? L.base + (lane < L.extra ? 1 : 0) : 0;
      const int p0 = lane * L.base + min(lane, L.extra);
      const long q0 = cidx(L, m, a, 0);
      T *row = sh + li * nsp;
      T lo[K], dg[K], up[K], rh[K];
#pragma unroll
      for (int j = 0; j < K; j++) {
        lo[j] = 0; dg[j] = 1; up[j] = 0; rh[j] = 0;
        if (j < cnt) {
          const int s = p0 + j;
          const long q = q0 + s;
          dg[j] = L.dg[q] - lmv * L.w[q];
          lo[j] = s > 0 ? -L.cs[q - 1] : T(0);
          up[j] = s < ns - 1 ? -L.cs[q] : T(0);
          rh[j] = row[lpad(s)];
          if (pin && s == pin_s) { dg[j] = 1; lo[j] = 0; up[j] = 0; rh[j] = 0; }
        }
      }
      line_elim<K, LPL, T, false>(
          ns, L.nch, cnt, true, [&](int j, T &l, T &d, T &u, T &r) { l = lo[j]; d = dg[j]; u = up[j]; r = rh[j]; },
          [&](int j, T v) { row[lpad(p0 + j)] = v; });
    }
  }
  __syncthreads();
  // phase 3: x of the lines, rows across threads
  for (int li = 0; li < NLW; li++) {
    const long line = line0 + li;
    if (line >= nl) break;
    const int a = 2 * (int)(line % nla) + col_a;
    const int lm_ = (int)(line / nla);
    const int m = L.coupled ? 2 * lm_ + col_m : lm_;
    if (done && done[group_of(L, m)]) continue;
    const long e0 = vidx(L, m, a, 0);
    for (int s = threadIdx.x; s < ns; s += 256) x[e0 + s] = sh[li * nsp + lpad(s)];
  }
}

// The LDS-staged relaxation with a workgroup's lines taken at one line
// position a over 256 / LPL consecutive modes (FFT path, where the level's
// coefficients do not depend on the mode): the line's coefficients are read
// from memory once per workgroup into LDS and shared by its modes, and each
// mode's right-hand side and x pass through LDS as in k_line_lds.  Each
// line goes to LPL lanes, which split it into min(LPL, ns / 2) chunks as the
// level's own chunking does for its lane count, and each lane reads its
// rows from LDS as the elimination asks for them (only the elimination's
// three coefficients per row stay in registers).  With the level's lane
// count the arithmetic is line_solve's, so the values are the same; with
// fewer lanes per line (longer chunks, a shorter interface system) the
// same line is solved with another elimination order.
template <int K, int LPL, typename T, bool FR, bool FULL = false>
__global__ __launch_bounds__(256) void k_line_modes(MGLevT<T> L, const T *__restrict__ lam, T *__restrict__ x,
                                                    const T *__restrict__ b, int col_a, int nla, int zero_nb,
                                                    int pin_s, const int *__restrict__ singular,
                                                    const int *__restrict__ done, const T *__restrict__ xc,
                                                    int cna) {
  constexpr int NLW = 256 / LPL;            // modes per workgroup
  constexpr int MB = NLW < 4 ? NLW : 4;     // modes whose loads are issued together
  extern __shared__ unsigned char smem_[];
  const int ns = L.ns, nsp = lpad_len(ns);
  // [nsp] each, rows at lpad: dg, w, cs, the couplings to a+1 and a-1
  T *cdg = reinterpret_cast<T *>(smem_);
  T *cw = cdg + nsp, *ccs = cw + nsp, *cap = ccs + nsp, *cam = cap + nsp;
  T *sh = cam + nsp;  // [NLW][nsp]: right-hand sides, then x
  // per mode slot of the workgroup: its mode, and -1 when skipped (past the
  // last mode, or converged), else the row to pin (ns: none); read once
  // from memory by one thread each
  __shared__ int sflag[NLW], smode[NLW];
  const int mb = (L.nm + NLW - 1) / NLW;   // mode batches per line position
  const int ai = blockIdx.x / mb, m0 = (blockIdx.x % mb) * NLW;  // the first mode slot
  // slots index the modes still iterating (the whole workgroup leaves when
  // its first slot is past them)
  const int nlive = L.mlist ? *L.mcount : L.nm;
  if (m0 >= nlive) return;
  const int a = 2 * ai + col_a;
  const long sa = ns, wrap = (long)(L.na - 1) * sa;
  const bool hp = a < L.na - 1 || L.per_a, hm = a > 0 || L.per_a;  // neighbour lines that exist
  const long dp = a < L.na - 1 ? sa : -wrap, dm = a > 0 ? -sa : wrap;  // their offsets
  const long q0 = cidx(L, 0, a, 0);
  if (threadIdx.x < NLW) {
    const int slot = m0 + threadIdx.x;
    const int m = slot < nlive ? (L.mlist ? L.mlist[slot] : slot) : -1;
    int f = -1;
    if (m >= 0 && !(done && done[m])) f = (pin_s >= 0 && singular && singular[m]) ? pin_s : ns;
    sflag[threadIdx.x] = f;
    smode[threadIdx.x] = m < 0 ? 0 : m;
  }
  for (int s = threadIdx.x; s < ns; s += 256) {
    const long q = q0 + s;
    const int o = lpad(s);
    cdg[o] = L.dg[q];
    cw[o] = L.w[q];
    ccs[o] = L.cs[q];
    cap[o] = hp ? L.ca[q] : T(0);
    cam[o] = a > 0 ? L.ca[q - sa] : (L.per_a ? L.ca[q + wrap] : T(0));
  }
  __syncthreads();
  // phase 1: right-hand sides of every mode of the workgroup, rows across
  // threads; MB modes' loads are issued together (no control flow between
  // them) so that their latencies overlap
  const int ra = cna != L.na;
  const int anp = a < L.na - 1 ? a + 1 : 0, anm = a > 0 ? a - 1 : L.na - 1;  // neighbour columns
  const int cnp = ra ? anp >> 1 : anp, cnm = ra ? anm >> 1 : anm;           // their coarse columns
  // per mode, base pointers of line a (and of the neighbour lines and their
  // coarse columns) indexed by the row: 32-bit offsets, no per-element
  // 64-bit address arithmetic
  const long mst = (long)L.na * ns, cmst = (long)cna * sa;  // elements per mode, fine and coarse
  const T *b0 = b + vidx(L, 0, a, 0), *x0 = x + vidx(L, 0, a, 0);
  for (int g0 = 0; g0 < NLW; g0 += MB) {
    bool live[MB];
    int pinrow[MB];
#pragma unroll
    for (int k = 0; k < MB; k++) {
      pinrow[k] = sflag[g0 + k];
      live[k] = pinrow[k] >= 0;
    }
    for (int s = threadIdx.x; s < ns; s += 256) {
      const unsigned us = (unsigned)s;
      const bool actp = xc && hp && L.act[q0 + s + dp], actm = xc && hm && L.act[q0 + s + dm];
      T bv[MB], xp[MB], xm[MB], cp[MB], cm[MB];
#pragma unroll
      for (int k = 0; k < MB; k++) {
        const long m = smode[g0 + k];
        const long mo = m * mst, co = m * cmst;
        bv[k] = live[k] ? (b0 + mo)[us] : T(0);
        xp[k] = live[k] && !zero_nb && hp ? (x0 + mo + dp)[us] : T(0);
        xm[k] = live[k] && !zero_nb && hm ? (x0 + mo + dm)[us] : T(0);
        cp[k] = live[k] && !zero_nb && actp ? (xc + co + (long)cnp * sa)[us] : T(0);
        cm[k] = live[k] && !zero_nb && actm ? (xc + co + (long)cnm * sa)[us] : T(0);
      }
      const int o = lpad(s);
#pragma unroll
      for (int k = 0; k < MB; k++) {
        if (!live[k]) continue;
        T r = bv[k];
        if (!zero_nb) {
          // nb_sum / nb_sum_pro with the coefficients from LDS, same order
          T nb = 0;
          if (hp) nb += cap[o] * (actp ? xp[k] + cp[k] : xp[k]);
          if (hm) nb += cam[o] * (actm ? xm[k] + cm[k] : xm[k]);
          r += nb;
        }
        if (s == pinrow[k]) r = 0;
        sh[(g0 + k) * nsp + o] = r;
      }
    }
  }
  __syncthreads();
  // phase 2: each lane segment eliminates its mode's line
  {
    const int li = threadIdx.x / LPL, m = smode[li];
    const int fl = sflag[li];
    if (fl >= 0) {  // a whole segment shares the test
      const T lmv = lam ? lam[m] : T(0);
      const bool pin = fl < ns;
      const int lane = (threadIdx.x & 63) & (LPL - 1);
      const int nch = FULL ? LPL : (ns >= 2 ? min(LPL, ns / 2) : 1);
      const int base = FULL ? K : (ns >= 2 ? ns / nch : 1), extra = FULL ? 0 : (ns >= 2 ? ns % nch : 0);
      const int cnt = lane < nch ? base + (lane < extra ? 1 : 0) : 0;
      const int p0 = lane * base + min(lane, extra);
      T *row = sh + li * nsp;
      line_elim<K, LPL, T, FR, FULL>(
          ns, nch, cnt, true,
          [&](int j, T &l, T &d, T &u, T &r) {
            const int s = p0 + j, o = lpad(s);
            d = cdg[o] - lmv * cw[o];
            l = s > 0 ? -ccs[lpad(s - 1)] : T(0);
            u = s < ns - 1 ? -ccs[o] : T(0);
            r = row[o];
            if (pin && s == pin_s) { d = 1; l = 0; u = 0; r = 0; }
          },
          [&](int j, T v) { row[lpad(p0 + j)] = v; });
    }
  }
  __syncthreads();
  // phase 3: x, rows across threads
  T *xo = x + vidx(L, 0, a, 0);
  for (int li = 0; li < NLW; li++) {
    if (sflag[li] < 0) continue;
    T *xl = xo + (long)smode[li] * mst;
    for (int s = threadIdx.x; s < ns; s += 256) xl[(unsigned)s] = sh[li * nsp + lpad(s)];
  }
}

// The coarse end of the V-cycle (levels t.l0 .. t.nl-1, the last one a
// single line per mode) for one mode per workgroup: zebra sweeps with one
// lane segment per line, residual restriction with the whole workgroup,
// the prolongation folded into the first post-smoothing sweep, barriers
// between the steps.  One launch replaces five per level.  FFT path only
// (independent modes).
// LX: the mode's x and b of every tail level live in LDS (b of the first
// level loaded at the start, x of the first level stored at the end), the
// level arrays then addressed as mode 0 (the coefficients do not depend on
// the mode); the arithmetic is the same, so the values are.
template <int K, int LPL, typename T, bool FR, bool LX, bool FULL = false>
__global__ __launch_bounds__(1024) void k_tail(MGTailT<T> t, const T *__restrict__ lam, int pin_s,
                                               const int *__restrict__ singular, const int *__restrict__ done) {
  extern __shared__ unsigned char smem_[];
  const int m = blockIdx.x;
  if (done && done[m]) return;
  const T lmv = lam[m];
  T *lds = reinterpret_cast<T *>(smem_);
  auto X = [&](int l) -> T * { return LX ? lds + t.ox[l] : t.x[l]; };
  auto B = [&](int l) -> T * { return LX ? lds + t.ob[l] : t.b[l]; };
  const int mi = LX ? 0 : m;  // the mode index the level arrays are addressed with
  if (LX) {
    const int n0 = t.lev[0].na * t.lev[0].ns;
    const T *bg = t.b[0] + vidx(t.lev[0], m, 0, 0);
    for (int e = threadIdx.x; e < n0; e += blockDim.x) lds[t.ob[0] + e] = bg[e];
    __syncthreads();
  }
  const int wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
  constexpr int SPW = 64 / LPL;                   // lines per wavefront
  const int hs = (threadIdx.x & 63) / LPL;        // this lane's segment
  const bool pin = pin_s >= 0 && singular && singular[m];
  // down: pre-smooth (colour 0 with zero neighbours, then colour 1), restrict
  for (int l = 0; l < t.nl - 1; l++) {
    const MGLevT<T> &L = t.lev[l];
    for (int c = 0; c < 2; c++) {
      const int nlc = (L.na - c + 1) / 2;  // lines a = 2 j + c of this colour
      for (int j0 = SPW * wv; j0 < nlc; j0 += SPW * nw) {
        const int j = j0 + hs;
        line_solve<K, LPL, T, FR, FULL>(L, lmv, X(l), B(l), mi, 2 * j + c, c == 0, -1, false, j < nlc);
      }
      __syncthreads();
    }
    const MGLevT<T> &Cl = t.lev[l + 1];
    const int ra = L.na != Cl.na;
    for (int e = threadIdx.x; e < Cl.na * Cl.ns; e += blockDim.x) {
      const int s = e % Cl.ns, I = e / Cl.ns;
      T sum = 0;
      for (int da = 0; da <= ra; da++) {
        const int a = ra ? 2 * I + da : I;
        if (a >= L.na || !L.act[cidx(L, mi, a, s)]) continue;
        sum += B(l)[vidx(L, mi, a, s)] - mx_at(L, lmv, X(l), mi, a, s);
      }
      B(l + 1)[vidx(Cl, mi, I, s)] = sum;
    }
    __syncthreads();
  }
  // coarsest: one line, solved exactly
  {
    const int l = t.nl - 1;
    if (wv == 0) line_solve<K, LPL, T, FR, FULL>(t.lev[l], lmv, X(l), B(l), mi, 0, 1, pin_s, pin, hs == 0);
    __syncthreads();
  }
  // up: post-smooth (colour 1, then colour 0); the colour-1 sweep reads its
  // colour-0 neighbours with the coarse correction added (nb_sum_pro), and
  // colour 0 is then relaxed from scratch, so the prolongation is never
  // stored (the values of a separate prolongation pass, bit for bit)
  for (int l = t.nl - 2; l >= 0; l--) {
    const MGLevT<T> &L = t.lev[l], &Cl = t.lev[l + 1];
    for (int c = 1; c >= 0; c--) {
      const int nlc = (L.na - c + 1) / 2;
      for (int j0 = SPW * wv; j0 < nlc; j0 += SPW * nw) {
        const int j = j0 + hs;
        line_solve<K, LPL, T, FR, FULL>(L, lmv, X(l), B(l), mi, 2 * j + c, 0, -1, false, j < nlc,
                                  c == 1 ? X(l + 1) : nullptr, Cl.na);
      }
      __syncthreads();
    }
  }
  if (LX) {  // the first level's x back to memory (the level above prolongs it)
    const int n0 = t.lev[0].na * t.lev[0].ns;
    T *xg = t.x[0] + vidx(t.lev[0], m, 0, 0);
    for (int e = threadIdx.x; e < n0; e += blockDim.x) xg[e] = lds[t.ox[0] + e];
  }
}

// coarse b = sum over the active fine members of (b - M x)
template <typename T>
__global__ __launch_bounds__(256) void k_restrict(MGLevT<T> F, MGLevT<T> Cl, const T *__restrict__ lamf,
                                                  const T *__restrict__ x, const T *__restrict__ b,
                                                  T *__restrict__ bc, const int *__restrict__ done) {
  const long idx = (long)blockIdx.x * 256 + threadIdx.x;
  const long n = (long)Cl.nm * Cl.na * Cl.ns;
  if (idx >= n) return;
  const int s = (int)(idx % Cl.ns);
  const long t = idx / Cl.ns;
  const int I = (int)(t % Cl.na), M = (int)(t / Cl.na);
  if (done && done[group_of(Cl, M)]) return;
  const int ra = F.na != Cl.na, rm = F.nm != Cl.nm;
  T sum = 0;
  for (int dm = 0; dm <= rm; dm++) {
    const int m = rm ? 2 * M + dm : M;
    if (m >= F.nm) break;
    const T lm = lamf ? lamf[m] : T(0);
    for (int da = 0; da <= ra; da++) {
      const int a = ra ? 2 * I + da : I;
      if (a >= F.na) break;
      if (!F.act[cidx(F, m, a, s)]) continue;
      sum += b[vidx(F, m, a, s)] - mx_at(F, lm, x, m, a, s);
    }
  }
  bc[idx] = sum;
}

// k_restrict on the FFT path with a workgroup's coarse cells at one coarse
// column I over RM consecutive modes: the coefficients of the fine columns
// 2I, 2I+1 (mode-independent there) are staged in LDS once and shared by the
// modes; rows across threads.  Same arithmetic as k_restrict (mx_at's
// order), so the same values.
template <typename T, int RM>
__global__ __launch_bounds__(256) void k_restrict_modes(MGLevT<T> F, MGLevT<T> Cl, const T *__restrict__ lamf,
                                                        const T *__restrict__ x, const T *__restrict__ b,
                                                        T *__restrict__ bc, const int *__restrict__ done) {
  extern __shared__ unsigned char smem_[];
  const int ns = F.ns;
  // per fine column (2): dg, w, cs, the couplings to a+1 and a-1, and act
  T *cdg = reinterpret_cast<T *>(smem_);
  T *cw = cdg + 2 * ns, *ccs = cw + 2 * ns, *cap = ccs + 2 * ns, *cam = cap + 2 * ns;
  unsigned char *cact = reinterpret_cast<unsigned char *>(cam + 2 * ns);
  const int ra = F.na != Cl.na;
  const int mb = (Cl.nm + RM - 1) / RM;
  const int I = blockIdx.x / mb, m0 = (blockIdx.x % mb) * RM;
  const int nda = ra ? 2 : 1;
  const long sa = ns, wrap = (long)(F.na - 1) * sa;
  // per mode of the workgroup, read once by one thread each: whether it is
  // relaxed (not past the last mode, not converged) and its eigenvalue
  // mode slots index the modes still iterating (the workgroup leaves when
  // its first slot is past them)
  __shared__ int slive[RM], smode[RM];
  __shared__ T slm[RM];
  const int nlive = F.mlist ? *F.mcount : Cl.nm;
  if (m0 >= nlive) return;
  if (threadIdx.x < RM) {
    const int slot = m0 + threadIdx.x;
    const int M = slot < nlive ? (F.mlist ? F.mlist[slot] : slot) : -1;
    const bool lv = M >= 0 && !(done && done[M]);
    slive[threadIdx.x] = lv;
    smode[threadIdx.x] = lv ? M : 0;
    slm[threadIdx.x] = lv && lamf ? lamf[M] : T(0);  // the FFT path: modes are not coarsened
  }
  for (int da = 0; da < nda; da++) {
    const int a = ra ? 2 * I + da : I;
    const bool ok = a < F.na;
    const long q0 = ok ? cidx(F, 0, a, 0) : 0;
    for (int s = threadIdx.x; s < ns; s += 256) {
      const int o = da * ns + s;
      if (!ok) { cact[o] = 0; continue; }
      const long q = q0 + s;
      cdg[o] = F.dg[q];
      cw[o] = F.w[q];
      ccs[o] = F.cs[q];
      cap[o] = (a < F.na - 1 || F.per_a) ? F.ca[q] : T(0);
      cam[o] = a > 0 ? F.ca[q - sa] : (F.per_a ? F.ca[q + wrap] : T(0));
      cact[o] = F.act[q];
    }
  }
  __syncthreads();
  // MG modes at a time, their loads issued together (no control flow between
  // them) so that their latencies overlap
  constexpr int MG = RM < 4 ? RM : 4;
  for (int g0 = 0; g0 < RM; g0 += MG) {
    bool live[MG];
    T lm[MG];
#pragma unroll
    for (int k = 0; k < MG; k++) {
      live[k] = slive[g0 + k];
      lm[k] = slm[g0 + k];
    }
    for (int s = threadIdx.x; s < ns; s += 256) {
      T sum[MG];
#pragma unroll
      for (int k = 0; k < MG; k++) sum[k] = 0;
      for (int da = 0; da < nda; da++) {
        const int a = ra ? 2 * I + da : I;
        if (a >= F.na) break;
        const int o = da * ns + s;
        if (!cact[o]) continue;
        const bool hp = a < F.na - 1 || F.per_a, hm = a > 0 || F.per_a;
        const long dp = a < F.na - 1 ? sa : -wrap, dm = a > 0 ? -sa : wrap;
        T xc[MG], xs0[MG], xs1[MG], xp[MG], xm[MG], bv[MG];
#pragma unroll
        for (int k = 0; k < MG; k++) {
          const long e = vidx(F, smode[g0 + k], a, s);
          xc[k] = live[k] ? x[e] : T(0);
          xs0[k] = live[k] && s > 0 ? x[e - 1] : T(0);
          xs1[k] = live[k] && s < ns - 1 ? x[e + 1] : T(0);
          xp[k] = live[k] && hp ? x[e + dp] : T(0);
          xm[k] = live[k] && hm ? x[e + dm] : T(0);
          bv[k] = live[k] ? b[e] : T(0);
        }
#pragma unroll
        for (int k = 0; k < MG; k++) {
          // mx_at: (dg - lm w) x - cs(s-1) x(s-1) - cs(s) x(s+1) - nb_sum
          T y = (cdg[o] - lm[k] * cw[o]) * xc[k];
          if (s > 0) y -= ccs[o - 1] * xs0[k];
          if (s < ns - 1) y -= ccs[o] * xs1[k];
          T r = 0;
          if (hp) r += cap[o] * xp[k];
          if (hm) r += cam[o] * xm[k];
          sum[k] += bv[k] - (y - r);
        }
      }
#pragma unroll
      for (int k = 0; k < MG; k++)
        if (live[k]) bc[vidx(Cl, smode[g0 + k], I, s)] = sum[k];
    }
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_prolong(MGLevT<T> F, MGLevT<T> Cl, const T *__restrict__ xc,
                                                 T *__restrict__ xf, const int *__restrict__ done) {
  const long idx = (long)blockIdx.x * 256 + threadIdx.x;
  const long n = (long)F.nm * F.na * F.ns;
  if (idx >= n) return;
  const int s = (int)(idx % F.ns);
  const long t = idx / F.ns;
  const int a = (int)(t % F.na), m = (int)(t / F.na);
  if (done && done[group_of(F, m)]) return;
  if (!F.act[cidx(F, m, a, s)]) return;
  const int I = F.na != Cl.na ? a >> 1 : a, M = F.nm != Cl.nm ? m >> 1 : m;
  xf[idx] += xc[vidx(Cl, M, I, s)];
}

// lam = base + shift in FP64 and FP32 (the Helmholtz plans' per-solve shift)
__global__ __launch_bounds__(256) void k_shift_lam(int n, const double *__restrict__ base, double shift,
                                                   double *__restrict__ l64, float *__restrict__ l32) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const double v = base[i] + shift;
  l64[i] = v;
  l32[i] = (float)v;
}

// a deterministic pattern in [-1, 1) (an integer hash of the index)
template <typename T>
__global__ __launch_bounds__(256) void k_pattern(long n, T *__restrict__ x) {
  const long idx = (long)blockIdx.x * 256 + threadIdx.x;
  if (idx >= n) return;
  unsigned h = (unsigned)idx * 2654435761u;
  h ^= h >> 15;
  h *= 2246822519u;
  h ^= h >> 13;
  x[idx] = (T)((double)(h >> 8) * (2.0 / 16777216.0) - 1.0);
}

template <typename T>
__global__ __launch_bounds__(256) void k_zero(long n, T *__restrict__ x) {
  const long idx = (long)blockIdx.x * 256 + threadIdx.x;
  if (idx < n) x[idx] = 0;
}

__global__ __launch_bounds__(256) void k_to_float(long n, const double *__restrict__ src, float *__restrict__ dst) {
  const long idx = (long)blockIdx.x * 256 + threadIdx.x;
  if (idx < n) dst[idx] = (float)src[idx];
}

__global__ __launch_bounds__(256) void k_izero(long n, int *__restrict__ x) {
  const long idx = (long)blockIdx.x * 256 + threadIdx.x;
  if (idx < n) x[idx] = 0;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// blockIdx.y = group, blockIdx.x = chunk of the group's elements
template <typename TY, typename TX2>
__global__ __launch_bounds__(256) void k_dot(long gsize, const double *__restrict__ x, const TY *__restrict__ y,
                                             const TX2 *__restrict__ x2, const double *__restrict__ y2,
                                             double *__restrict__ part, int nchunk, const int *__restrict__ done,
                                             const int *__restrict__ only) {
  __shared__ double sh[2][4];
  const int g = blockIdx.y;
  const bool skip = (done && done[g]) || (only && !only[g]);
  double s1 = 0.0, s2 = 0.0;
  if (!skip) {
    const long base = (long)g * gsize;
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < gsize; i += (long)nchunk * 256) {
      if (x) s1 += x[base + i] * (double)y[base + i];
      if (x2) s2 += y2 ? (double)x2[base + i] * y2[base + i] : (double)x2[base + i];
    }
  }
  s1 = wave_sum(s1);
  s2 = wave_sum(s2);
  if ((threadIdx.x & 63) == 0) { sh[0][threadIdx.x >> 6] = s1; sh[1][threadIdx.x >> 6] = s2; }
  __syncthreads();
  if (threadIdx.x == 0) {
    part[((long)g * nchunk + blockIdx.x) * 2] = sh[0][0] + sh[0][1] + sh[0][2] + sh[0][3];
    part[((long)g * nchunk + blockIdx.x) * 2 + 1] = sh[1][0] + sh[1][1] + sh[1][2] + sh[1][3];
  }
}

// the CG's operator application and its first dot product in one pass:
// q = M p and the per-group partial sums of p.q, laid out as k_dot's (same
// element order per thread, so the sums are k_dot's to the bit)
__global__ __launch_bounds__(256) void k_apply_dot(MGLev L, const double *__restrict__ lam,
                                                   const double *__restrict__ p, double *__restrict__ q,
                                                   double *__restrict__ part, int nchunk, long gsize,
                                                   const int *__restrict__ done) {
  __shared__ double sh[4];
  const int g = blockIdx.y;
  double s1 = 0.0;
  if (!(done && done[g])) {
    const long base = (long)g * gsize;
    const long i0 = (long)blockIdx.x * 256 + threadIdx.x, step = (long)nchunk * 256;
    int m, a, s;
    pos_of(L, base + i0, m, a, s);
    const int da = (int)(step / L.ns), ds = (int)(step % L.ns);
    for (long i = i0; i < gsize; i += step) {
      const long idx = base + i;
      const double y = mx_at(L, lam ? lam[m] : 0.0, p, m, a, s);
      q[idx] = y;
      s1 += p[idx] * y;
      pos_step(L, da, ds, m, a, s);
    }
  }
  s1 = wave_sum(s1);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s1;
  __syncthreads();
  if (threadIdx.x == 0) {
    part[((long)g * nchunk + blockIdx.x) * 2] = sh[0] + sh[1] + sh[2] + sh[3];
    part[((long)g * nchunk + blockIdx.x) * 2 + 1] = 0.0;
  }
}

// x += alpha p ; r -= alpha q (alpha = rz / pq per group) with the partial
// sums of r.r and, when zf is given, zf.r (the flexible-CG cross term), and
// the FP32 copy of r when rf is given
__global__ __launch_bounds__(256) void k_cg_update_dot(long gsize, const double *__restrict__ rz,
                                                       const double *__restrict__ pq, double *__restrict__ x,
                                                       const double *__restrict__ p, double *__restrict__ r,
                                                       const double *__restrict__ q, float *__restrict__ rf,
                                                       const float *__restrict__ zf, double *__restrict__ part,
                                                       int nchunk, const int *__restrict__ done) {
  __shared__ double sh[2][4];
  const int g = blockIdx.y;
  double s1 = 0.0, s2 = 0.0;
  if (!done[g]) {
    const double den = pq[g * 2];
    // p.Mp > 0 for an SPD operator and p != 0; anything else (p = 0, or a
    // preconditioner output that is not usable) skips the step instead of
    // spreading an overflow or a NaN
    double al = den > 0.0 ? rz[g * 2] / den : 0.0;
    if (!(__builtin_fabs(al) < 1e300)) al = 0.0;
    const long base = (long)g * gsize;
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < gsize; i += (long)nchunk * 256) {
      const long idx = base + i;
      if (al != 0.0) x[idx] += al * p[idx];  // a skipped step leaves x and r as they are
      const double rn = al != 0.0 ? r[idx] - al * q[idx] : r[idx];
      r[idx] = rn;
      s1 += rn * rn;
      if (rf) rf[idx] = (float)rn;
      if (zf) s2 += (double)zf[idx] * rn;
    }
  }
  s1 = wave_sum(s1);
  s2 = wave_sum(s2);
  if ((threadIdx.x & 63) == 0) { sh[0][threadIdx.x >> 6] = s1; sh[1][threadIdx.x >> 6] = s2; }
  __syncthreads();
  if (threadIdx.x == 0) {
    part[((long)g * nchunk + blockIdx.x) * 2] = sh[0][0] + sh[0][1] + sh[0][2] + sh[0][3];
    part[((long)g * nchunk + blockIdx.x) * 2 + 1] = sh[1][0] + sh[1][1] + sh[1][2] + sh[1][3];
  }
}

// warm start of a solve from x0 (the previous solve's solution, in x): with
// q = M x0, r = b - q (r holds b) with its r.r and rf = FP32 r, on the
// groups whose b is not zero; a group with b = 0 has the solution 0 (x and
// rf cleared, r is already 0)
__global__ __launch_bounds__(256) void k_warm_residual(long gsize, const double *__restrict__ bb,
                                                       double *__restrict__ x, double *__restrict__ r,
                                                       const double *__restrict__ q, float *__restrict__ rf,
                                                       double *__restrict__ part, int nchunk) {
  __shared__ double sh[4];
  const int g = blockIdx.y;
  const bool zero = !(bb[g * 2] > 0.0);
  const long base = (long)g * gsize;
  double s1 = 0.0;
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < gsize; i += (long)nchunk * 256) {
    const long idx = base + i;
    if (zero) {
      x[idx] = 0.0;
      if (rf) rf[idx] = 0.0f;
      continue;
    }
    const double rn = r[idx] - q[idx];
    r[idx] = rn;
    s1 += rn * rn;
    if (rf) rf[idx] = (float)rn;
  }
  s1 = wave_sum(s1);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s1;
  __syncthreads();
  if (threadIdx.x == 0) {
    part[((long)g * nchunk + blockIdx.x) * 2] = sh[0] + sh[1] + sh[2] + sh[3];
    part[((long)g * nchunk + blockIdx.x) * 2 + 1] = 0.0;
  }
}

// (M v)(m, a, s) of an FP64 level where v is given element-wise by f(e)
template <typename F>
__device__ __forceinline__ double mx_f(const MGLev &L, double lm, F &&f, int m, int a, int s, long e, long q) {
  double y = (L.dg[q] - lm * L.w[q]) * f(e);
  if (s > 0) y -= L.cs[q - 1] * f(e - 1);
  if (s < L.ns - 1) y -= L.cs[q] * f(e + 1);
  double r = 0;
  const long sa = L.ns;
  if (a < L.na - 1) r += L.ca[q] * f(e + sa);
  else if (L.per_a) r += L.ca[q] * f(e - (long)(L.na - 1) * sa);
  if (a > 0) r += L.ca[q - sa] * f(e - sa);
  else if (L.per_a) r += L.ca[q + (long)(L.na - 1) * sa] * f(e + (long)(L.na - 1) * sa);
  if (L.coupled) {
    const long sm = (long)L.na * L.ns;
    if (m < L.nm - 1) r += L.cm[q] * f(e + sm);
    else if (L.per_m) r += L.cm[q] * f(e - (long)(L.nm - 1) * sm);
    if (m > 0) r += L.cm[q - L.qm] * f(e - sm);
    else if (L.per_m) r += L.cm[q + (long)(L.nm - 1) * L.qm] * f(e + (long)(L.nm - 1) * sm);
  }
  return y - r;
}

// elements per round of a thread in the fused CG kernels
constexpr int CG_U = 4;

// the operator's coefficients at one row s of a level whose coefficients do
// not depend on the position a (auni): read once by a thread whose elements
// all lie on row s
struct MxRow {
  double dg, w, csm, csp, ca;
};
__device__ __forceinline__ MxRow mx_row(const MGLev &L, int s) {
  MxRow c;
  c.dg = L.dg[s];
  c.w = L.w[s];
  c.csm = s > 0 ? L.cs[s - 1] : 0.0;
  c.csp = s < L.ns - 1 ? L.cs[s] : 0.0;
  c.ca = L.ca[s];
  return c;
}
// mx_f on such a level from the row's coefficients: the same values in the
// same order (ca at a and at a - 1 are the same number there)
template <typename F>
__device__ __forceinline__ double mx_r(const MGLev &L, double lm, F &&f, int a, int s, long e, const MxRow &c) {
  double y = (c.dg - lm * c.w) * f(e);
  if (s > 0) y -= c.csm * f(e - 1);
  if (s < L.ns - 1) y -= c.csp * f(e + 1);
  double r = 0;
  const long sa = L.ns;
  if (a < L.na - 1) r += c.ca * f(e + sa);
  else if (L.per_a) r += c.ca * f(e - (long)(L.na - 1) * sa);
  if (a > 0) r += c.ca * f(e - sa);
  else if (L.per_a) r += c.ca * f(e + (long)(L.na - 1) * sa);
  return y - r;
}

__device__ __forceinline__ double cg_beta(int g, const double *__restrict__ rzn, const double *__restrict__ rzo,
                                          const double *__restrict__ cross) {
  if (!rzo) return 0.0;
  const double den = rzo[g * 2];
  const double num = rzn[g * 2] - (cross ? cross[g * 2 + 1] : 0.0);
  return den != 0.0 ? num / den : 0.0;
}

// The search direction and the operator in one pass: p = z + beta p_prev
// (rounded to TP, the stored type), and the per-group partial sums of
// p . (M p), M p computed from the new directions of the neighbours formed
// on the fly (p_prev and p are different buffers).
// STAGE (2D levels): the workgroup's chunk of new directions with one line
// before and after it is formed once into LDS (dynamic, (chunk + 2 ns) x
// TP) and the operator reads it there; only the wrap of a periodic a at the
// group's ends forms its neighbour again.  Same values either way.
template <typename TZ, typename TP, bool STAGE = false>
__global__ __launch_bounds__(256) void k_dir_apply_dot(MGLev L, const double *__restrict__ lam, long gsize,
                                                       const double *__restrict__ rzn, const double *__restrict__ rzo,
                                                       const double *__restrict__ cross, const TZ *__restrict__ z,
                                                       const TP *__restrict__ pp, TP *__restrict__ p,
                                                       double *__restrict__ part, int nchunk,
                                                       const int *__restrict__ done) {
  extern __shared__ unsigned char dyn_smem[];
  __shared__ double sh[4];
  const int g = blockIdx.y;
  double s1 = 0.0;
  if (!done[g]) {
    const double be = cg_beta(g, rzn, rzo, cross);
    auto pnew = [&](long e) -> double {
      if (be == 0.0) return (double)(TP)(double)z[e];
      return (double)(TP)((double)z[e] + be * (double)pp[e]);
    };
    const long base = (long)g * gsize;
    // a contiguous chunk of the group per workgroup: the neighbour lines the
    // operator reads are mostly this workgroup's own (cache hits)
    const long csz = ((gsize + nchunk - 1) / nchunk + 255) / 256 * 256;
    const long iend = min(gsize, ((long)blockIdx.x + 1) * csz);
    const long i0 = (long)blockIdx.x * csz + threadIdx.x;
    TP *sp = reinterpret_cast<TP *>(dyn_smem);
    const long k0s = STAGE ? max(0L, (long)blockIdx.x * csz - L.ns) : 0;
    const long k1s = STAGE ? min(gsize, iend + L.ns) : 0;
    if (STAGE) {
      for (long k = k0s + threadIdx.x; k < k1s; k += 256) sp[k - k0s] = (TP)pnew(base + k);
      __syncthreads();
    }
    auto pn = [&](long e) -> double {
      if (STAGE) {
        const long k = e - base;
        if (k >= k0s && k < k1s) return (double)sp[k - k0s];
      }
      return pnew(e);
    };
    int m, a, s;
    pos_of(L, base + i0, m, a, s);
    const int da = 256 / L.ns, ds = 256 % L.ns;
    // a thread's elements keep their row s when 256 is a multiple of ns: on
    // a level with position-independent coefficients they are read once
    const bool rowc = L.auni && ds == 0;
    const MxRow c = rowc ? mx_row(L, s) : MxRow{};
    // CG_U elements per round, their loads issued before any store so that
    // their latencies overlap; the sum keeps the element order
    for (long i = i0; i < iend; i += 256 * CG_U) {
      double pc[CG_U], y[CG_U];
#pragma unroll
      for (int u = 0; u < CG_U; u++) {
        const long idx = base + i + 256 * u;
        const bool ok = i + 256 * u < iend;
        const double lm = lam ? lam[m] : 0.0;
        pc[u] = ok ? pn(idx) : 0.0;
        y[u] = !ok ? 0.0 : rowc ? mx_r(L, lm, pn, a, s, idx, c) : mx_f(L, lm, pn, m, a, s, idx, cidx(L, m, a, s));
        pos_step(L, da, ds, m, a, s);
      }
#pragma unroll
      for (int u = 0; u < CG_U; u++) {
        if (i + 256 * u >= iend) break;
        p[base + i + 256 * u] = (TP)pc[u];
        s1 += pc[u] * y[u];
      }
    }
  }
  s1 = wave_sum(s1);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s1;
  __syncthreads();
  if (threadIdx.x == 0) {
    part[((long)g * nchunk + blockIdx.x) * 2] = sh[0] + sh[1] + sh[2] + sh[3];
    part[((long)g * nchunk + blockIdx.x) * 2 + 1] = 0.0;
  }
}

// k_dir_apply_dot on a 2D level with position-independent coefficients
// whose rows divide 256 (every thread keeps its row s): the workgroup's
// chunk is whole lines a0..a1-1, and its new directions, with those of the
// neighbour lines a0-1 and a1 (wrapped on a periodic a), are formed once
// into LDS; M p then reads them there instead of forming each element's
// four neighbours again from z and p_prev.  Same values, same order of the
// sums, as k_dir_apply_dot.
template <typename TZ, typename TP>
__global__ __launch_bounds__(256) void k_dir_apply_dot_lds(MGLev L, const double *__restrict__ lam, long gsize,
                                                           const double *__restrict__ rzn,
                                                           const double *__restrict__ rzo,
                                                           const double *__restrict__ cross,
                                                           const TZ *__restrict__ z, const TP *__restrict__ pp,
                                                           TP *__restrict__ p, double *__restrict__ part, int nchunk,
                                                           const int *__restrict__ done) {
  extern __shared__ unsigned char dyn_smem[];
  TP *sp = reinterpret_cast<TP *>(dyn_smem);  // (lines + 2) x ns: below, own lines, above
  __shared__ double sh[4];
  const int g = blockIdx.y;
  double s1 = 0.0;
  if (!done[g]) {
    const double be = cg_beta(g, rzn, rzo, cross);
    auto pn = [&](long e) -> TP {
      if (be == 0.0) return (TP)(double)z[e];
      return (TP)((double)z[e] + be * (double)pp[e]);
    };
    const int ns = L.ns, na = L.na;
    const long base = (long)g * gsize;
    const long csz = ((gsize + nchunk - 1) / nchunk + 255) / 256 * 256;
    const long c0 = (long)blockIdx.x * csz, c1 = min(gsize, c0 + csz);
    if (c0 < c1) {
      const int a0 = (int)(c0 / ns), nl = (int)((c1 - c0) / ns);
      // the neighbour lines below and above the chunk (-1: none)
      const int lo = a0 > 0 ? a0 - 1 : (L.per_a ? na - 1 : -1);
      const int hi = a0 + nl < na ? a0 + nl : (L.per_a ? 0 : -1);
      const int tot = (nl + 2) * ns;
      for (int k = threadIdx.x; k < tot; k += 256) {
        const int la = k / ns, s = k - la * ns;
        const int a = la == 0 ? lo : (la == nl + 1 ? hi : a0 + la - 1);
        if (a >= 0) sp[k] = pn(base + (long)a * ns + s);
      }
      __syncthreads();
      const int s = threadIdx.x % ns;
      const MxRow c = mx_row(L, s);
      const double lm = lam ? lam[g] : 0.0;
      for (int k = threadIdx.x; k < nl * ns; k += 256) {
        const int la = k / ns + 1, a = a0 + la - 1;
        const TP *row = sp + la * ns;
        const double pc = (double)row[s];
        double y = (c.dg - lm * c.w) * pc;
        if (s > 0) y -= c.csm * (double)row[s - 1];
        if (s < ns - 1) y -= c.csp * (double)row[s + 1];
        double r = 0;
        if (a < na - 1 || L.per_a) r += c.ca * (double)row[s + ns];
        if (a > 0 || L.per_a) r += c.ca * (double)row[s - ns];
        y = y - r;
        p[base + c0 + k] = row[s];
        s1 += pc * y;
      }
    }
  }
  s1 = wave_sum(s1);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s1;
  __syncthreads();
  if (threadIdx.x == 0) {
    part[((long)g * nchunk + blockIdx.x) * 2] = sh[0] + sh[1] + sh[2] + sh[3];
    part[((long)g * nchunk + blockIdx.x) * 2 + 1] = 0.0;
  }
}

// x += alpha p ; r -= alpha (M p) with alpha = rz / pq per group, M p
// recomputed from p (bit for bit the value k_dir_apply_dot summed), the
// partial sums of r.r and, when zf is given, zf.r, and rf = (float) r when
// given
// STAGE: as k_dir_apply_dot's, the chunk's directions with one line before
// and after it in LDS
template <typename TP, bool STAGE = false>
__global__ __launch_bounds__(256) void k_update_apply_dot(MGLev L, const double *__restrict__ lam, long gsize,
                                                          const double *__restrict__ rz,
                                                          const double *__restrict__ pq, double *__restrict__ x,
                                                          const TP *__restrict__ p, double *__restrict__ r,
                                                          float *__restrict__ rf, const float *__restrict__ zf,
                                                          double *__restrict__ part, int nchunk,
                                                          const int *__restrict__ done, int first) {
  extern __shared__ unsigned char dyn_smem[];
  __shared__ double sh[2][4];
  const int g = blockIdx.y;
  double s1 = 0.0, s2 = 0.0;
  if (!done[g]) {
    const double den = pq[g * 2];
    // p.Mp > 0 for an SPD operator and p != 0; anything else (p = 0, or a
    // preconditioner output that is not usable) skips the step instead of
    // spreading an overflow or a NaN
    double al = den > 0.0 ? rz[g * 2] / den : 0.0;
    if (!(__builtin_fabs(al) < 1e300)) al = 0.0;
    const long base = (long)g * gsize;
    // a contiguous chunk of the group per workgroup: the neighbour lines the
    // operator reads are mostly this workgroup's own (cache hits)
    const long csz = ((gsize + nchunk - 1) / nchunk + 255) / 256 * 256;
    const long iend = min(gsize, ((long)blockIdx.x + 1) * csz);
    const long i0 = (long)blockIdx.x * csz + threadIdx.x;
    TP *sp = reinterpret_cast<TP *>(dyn_smem);
    const long k0s = STAGE ? max(0L, (long)blockIdx.x * csz - L.ns) : 0;
    const long k1s = STAGE ? min(gsize, iend + L.ns) : 0;
    if (STAGE) {
      for (long k = k0s + threadIdx.x; k < k1s; k += 256) sp[k - k0s] = p[base + k];
      __syncthreads();
    }
    auto pv = [&](long e) -> double {
      if (STAGE) {
        const long k = e - base;
        if (k >= k0s && k < k1s) return (double)sp[k - k0s];
      }
      return (double)p[e];
    };
    int m, a, s;
    pos_of(L, base + i0, m, a, s);
    const int da = 256 / L.ns, ds = 256 % L.ns;
    // position-independent coefficients on a fixed row: read once (as in
    // k_dir_apply_dot, so the same M p)
    const bool rowc = L.auni && ds == 0;
    const MxRow c = rowc ? mx_row(L, s) : MxRow{};
    // CG_U elements per round, their loads issued before any store so that
    // their latencies overlap; the sums keep the element order
    for (long i = i0; i < iend; i += 256 * CG_U) {
      double q[CG_U], pc[CG_U], xv[CG_U], rv[CG_U], zv[CG_U];
#pragma unroll
      for (int u = 0; u < CG_U; u++) {
        const long idx = base + i + 256 * u;
        const bool ok = i + 256 * u < iend;
        const double lm = lam ? lam[m] : 0.0;
        q[u] = !ok ? 0.0 : rowc ? mx_r(L, lm, pv, a, s, idx, c) : mx_f(L, lm, pv, m, a, s, idx, cidx(L, m, a, s));
        pc[u] = ok ? pv(idx) : 0.0;
        // x and r are read and written once here and not again until after
        // the next V-cycle: streamed past the caches
        xv[u] = ok && !first ? __builtin_nontemporal_load(&x[idx]) : 0.0;
        rv[u] = ok ? __builtin_nontemporal_load(&r[idx]) : 0.0;
        zv[u] = ok && zf ? (double)zf[idx] : 0.0;
        pos_step(L, da, ds, m, a, s);
      }
#pragma unroll
      for (int u = 0; u < CG_U; u++) {
        if (i + 256 * u >= iend) break;
        const long idx = base + i + 256 * u;
        // a skipped step (al = 0) leaves x and r as they are, also where a
        // broken direction holds an overflow or a NaN (0 * inf is a NaN)
        __builtin_nontemporal_store(al != 0.0 ? xv[u] + al * pc[u] : xv[u], &x[idx]);
        const double rn = al != 0.0 ? rv[u] - al * q[u] : rv[u];
        __builtin_nontemporal_store(rn, &r[idx]);
        s1 += rn * rn;
        if (rf) rf[idx] = (float)rn;
        if (zf) s2 += zv[u] * rn;
      }
    }
  } else if (first) {  // the solve's first step: a converged group's solution is 0
    const long csz = ((gsize + nchunk - 1) / nchunk + 255) / 256 * 256;
    const long c0 = (long)blockIdx.x * csz, c1 = min(gsize, c0 + csz);
    for (long i = c0 + threadIdx.x; i < c1; i += 256) x[(long)g * gsize + i] = 0.0;
  }
  s1 = wave_sum(s1);
  s2 = wave_sum(s2);
  if ((threadIdx.x & 63) == 0) { sh[0][threadIdx.x >> 6] = s1; sh[1][threadIdx.x >> 6] = s2; }
  __syncthreads();
  if (threadIdx.x == 0) {
    part[((long)g * nchunk + blockIdx.x) * 2] = sh[0][0] + sh[0][1] + sh[0][2] + sh[0][3];
    part[((long)g * nchunk + blockIdx.x) * 2 + 1] = sh[1][0] + sh[1][1] + sh[1][2] + sh[1][3];
  }
}

// k_update_apply_dot on the levels k_dir_apply_dot_lds serves: the chunk's
// directions with the two neighbour lines are staged in LDS once, and M p
// reads them there; same values, same order of the sums
template <typename TP>
__global__ __launch_bounds__(256) void k_update_apply_dot_lds(MGLev L, const double *__restrict__ lam, long gsize,
                                                              const double *__restrict__ rz,
                                                              const double *__restrict__ pq, double *__restrict__ x,
                                                              const TP *__restrict__ p, double *__restrict__ r,
                                                              float *__restrict__ rf, const float *__restrict__ zf,
                                                              double *__restrict__ part, int nchunk,
                                                              const int *__restrict__ done, int first) {
  extern __shared__ unsigned char dyn_smem[];
  TP *sp = reinterpret_cast<TP *>(dyn_smem);  // (lines + 2) x ns: below, own lines, above
  __shared__ double sh[2][4];
  const int g = blockIdx.y;
  double s1 = 0.0, s2 = 0.0;
  if (!done[g]) {
    const double den = pq[g * 2];
    // as k_update_apply_dot: an unusable step length skips the step
    double al = den > 0.0 ? rz[g * 2] / den : 0.0;
    if (!(__builtin_fabs(al) < 1e300)) al = 0.0;
    const int ns = L.ns, na = L.na;
    const long base = (long)g * gsize;
    const long csz = ((gsize + nchunk - 1) / nchunk + 255) / 256 * 256;
    const long c0 = (long)blockIdx.x * csz, c1 = min(gsize, c0 + csz);
    if (c0 < c1) {
      const int a0 = (int)(c0 / ns), nl = (int)((c1 - c0) / ns);
      const int lo = a0 > 0 ? a0 - 1 : (L.per_a ? na - 1 : -1);
      const int hi = a0 + nl < na ? a0 + nl : (L.per_a ? 0 : -1);
      const int tot = (nl + 2) * ns;
      for (int k = threadIdx.x; k < tot; k += 256) {
        const int la = k / ns, s = k - la * ns;
        const int a = la == 0 ? lo : (la == nl + 1 ? hi : a0 + la - 1);
        if (a >= 0) sp[k] = p[base + (long)a * ns + s];
      }
      __syncthreads();
      const int s = threadIdx.x % ns;
      const MxRow c = mx_row(L, s);
      const double lm = lam ? lam[g] : 0.0;
      const int nel = nl * ns;
      for (int k0 = threadIdx.x; k0 < nel; k0 += 256 * CG_U) {
        double q[CG_U], pc[CG_U], xv[CG_U], rv[CG_U], zv[CG_U];
#pragma unroll
        for (int u = 0; u < CG_U; u++) {
          const int k = k0 + 256 * u;
          const bool ok = k < nel;
          const long idx = base + c0 + k;
          const int la = ok ? k / ns + 1 : 1, a = a0 + la - 1;
          const TP *row = sp + la * ns;
          double y = 0.0;
          if (ok) {
            y = (c.dg - lm * c.w) * (double)row[s];
            if (s > 0) y -= c.csm * (double)row[s - 1];
            if (s < ns - 1) y -= c.csp * (double)row[s + 1];
            double rr = 0;
            if (a < na - 1 || L.per_a) rr += c.ca * (double)row[s + ns];
            if (a > 0 || L.per_a) rr += c.ca * (double)row[s - ns];
            y = y - rr;
          }
          q[u] = y;
          pc[u] = ok ? (double)row[s] : 0.0;
          xv[u] = ok && !first ? __builtin_nontemporal_load(&x[idx]) : 0.0;
          rv[u] = ok ? __builtin_nontemporal_load(&r[idx]) : 0.0;
          zv[u] = ok && zf ? (double)zf[idx] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < CG_U; u++) {
          const int k = k0 + 256 * u;
          if (k >= nel) break;
          const long idx = base + c0 + k;
          __builtin_nontemporal_store(al != 0.0 ? xv[u] + al * pc[u] : xv[u], &x[idx]);
          const double rn = al != 0.0 ? rv[u] - al * q[u] : rv[u];
          __builtin_nontemporal_store(rn, &r[idx]);
          s1 += rn * rn;
          if (rf) rf[idx] = (float)rn;
          if (zf) s2 += zv[u] * rn;
        }
      }
    }
  } else if (first) {  // the solve's first step: a converged group's solution is 0
    const long csz = ((gsize + nchunk - 1) / nchunk + 255) / 256 * 256;
    const long c0 = (long)blockIdx.x * csz, c1 = min(gsize, c0 + csz);
    for (long i = c0 + threadIdx.x; i < c1; i += 256) x[(long)g * gsize + i] = 0.0;
  }
  s1 = wave_sum(s1);
  s2 = wave_sum(s2);
  if ((threadIdx.x & 63) == 0) { sh[0][threadIdx.x >> 6] = s1; sh[1][threadIdx.x >> 6] = s2; }
  __syncthreads();
  if (threadIdx.x == 0) {
    part[((long)g * nchunk + blockIdx.x) * 2] = sh[0][0] + sh[0][1] + sh[0][2] + sh[0][3];
    part[((long)g * nchunk + blockIdx.x) * 2 + 1] = sh[1][0] + sh[1][1] + sh[1][2] + sh[1][3];
  }
}

// one wavefront per group sums its chunks (skipped groups keep their values)
__global__ __launch_bounds__(64) void k_dot_final(const double *__restrict__ part, int nchunk, double *__restrict__ out,
                                                  const int *__restrict__ done, const int *__restrict__ only) {
  const int g = blockIdx.x;
  if ((done && done[g]) || (only && !only[g])) return;
  double a = 0.0, b = 0.0;
  for (int c = threadIdx.x; c < nchunk; c += 64) {
    a += part[((long)g * nchunk + c) * 2];
    b += part[((long)g * nchunk + c) * 2 + 1];
  }
  a = wave_sum(a);
  b = wave_sum(b);
  if (threadIdx.x == 0) { out[g * 2] = a; out[g * 2 + 1] = b; }
}

template <typename T>
__global__ __launch_bounds__(256) void k_cg_dir(int ng, long gsize, const double *__restrict__ rzn,
                                                const double *__restrict__ rzo, const double *__restrict__ cross,
                                                const T *__restrict__ z, double *__restrict__ p,
                                                const int *__restrict__ done) {
  const long idx = (long)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (long)ng * gsize) return;
  const int g = (int)(idx / gsize);
  if (done[g]) return;
  double be = 0.0;
  if (rzo) {
    const double den = rzo[g * 2];
    const double num = rzn[g * 2] - (cross ? cross[g * 2 + 1] : 0.0);
    be = den != 0.0 ? num / den : 0.0;
  }
  p[idx] = (double)z[idx] + be * p[idx];
}

// over the singular groups only (glist, one per blockIdx.y): the grid covers
// their gsize elements each, not the whole vector
template <typename T>
__global__ __launch_bounds__(256) void k_sub_mean(MGLev L, T *__restrict__ x, const double *__restrict__ sums,
                                                  const int *__restrict__ glist, long gsize,
                                                  const double *__restrict__ nact, const int *__restrict__ done) {
  const int g = glist[blockIdx.y];
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= gsize || (done && done[g])) return;
  const long idx = (long)g * gsize + i;
  int m, a, s;
  pos_of(L, idx, m, a, s);
  if (!L.act[cidx(L, m, a, s)]) return;
  x[idx] -= (T)(sums[g * 2 + 1] / nact[g]);
}

// one workgroup: updates the done flags, counts the groups left and (list)
// writes them in increasing order, 256 groups per round (ballot + popcount
// positions within a wavefront, wavefront totals through LDS)
__global__ __launch_bounds__(256) void k_flags(int ng, const double *__restrict__ rr, const double *__restrict__ bb,
                                               double tol2, int *__restrict__ done, int *__restrict__ left,
                                               int *__restrict__ list, int it) {
  __shared__ int wsum[4];
  __shared__ int base;
  if (threadIdx.x == 0) base = 0;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int g0 = 0; g0 < ng; g0 += 256) {
    const int g = g0 + threadIdx.x;
    int act = 0;
    if (g < ng) {
      if (!done[g] && rr[g * 2] <= tol2 * bb[g * 2]) done[g] = 1;
      act = !done[g];
    }
    const unsigned long long bal = __ballot(act);
    const int pre = __popcll(bal & ((1ull << lane) - 1ull));
    if (lane == 0) wsum[wv] = __popcll(bal);
    __syncthreads();
    int off = base;
    for (int k = 0; k < wv; k++) off += wsum[k];
    const int tot = wsum[0] + wsum[1] + wsum[2] + wsum[3];
    if (act && list) list[off + pre] = g;
    __syncthreads();  // base and wsum read by every thread before they change
    if (threadIdx.x == 0) base += tot;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    left[0] = base;
    if (base == 0 && left[1] == 0) left[1] = it + 1;
  }
}

// strided 3D copy; TILE 1: lanes along a on the read side (src stride(a) 1),
// along s on the write side (dst stride(s) 1); TILE 2 the reverse
constexpr int TT = 32;
template <int TILE>
__global__ __launch_bounds__(256) void k_copy3(const double *__restrict__ src, long sm, long sa, long ss,
                                               double *__restrict__ dst, long dm, long da, long ds, int ea, int es,
                                               double scale) {
  const long m = blockIdx.z;
  if (TILE == 0) {
    const int s = blockIdx.x * 256 + threadIdx.x, a = blockIdx.y;
    if (s >= es) return;
    dst[m * dm + a * da + s * ds] = scale * src[m * sm + a * sa + s * ss];
    return;
  }
  __shared__ double t[TT][TT + 1];
  const int a0 = blockIdx.y * TT, s0 = blockIdx.x * TT;
  const int l = threadIdx.x % TT, r = threadIdx.x / TT;  // 8 rows per pass
  if (TILE == 1) {
#pragma unroll
    for (int k = 0; k < TT; k += 8) {
      const int a = a0 + l, s = s0 + r + k;
      if (a < ea && s < es) t[r + k][l] = src[m * sm + a * sa + s * ss];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < TT; k += 8) {
      const int s = s0 + l, a = a0 + r + k;
      if (a < ea && s < es) dst[m * dm + a * da + s * ds] = scale * t[l][r + k];
    }
  } else {
#pragma unroll
    for (int k = 0; k < TT; k += 8) {
      const int s = s0 + l, a = a0 + r + k;
      if (a < ea && s < es) t[r + k][l] = src[m * sm + a * sa + s * ss];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < TT; k += 8) {
      const int a = a0 + l, s = s0 + r + k;
      if (a < ea && s < es) dst[m * dm + a * da + s * ds] = scale * t[l][r + k];
    }
  }
}

unsigned blocks_for(long n) { return (unsigned)((n + 255) / 256); }

}  // namespace

namespace {
// 0: rows read per lane; 1: right-hand sides staged through LDS; 2 (the
// default): 1 with the lines of a workgroup at one position over several
// modes (FFT path; elsewhere 1)
int g_line_variant = -1;  // forced variant, or -1: each level's own
// the eliminations of the mode-batched sweeps and of the fused tail with
// refined hardware reciprocals instead of IEEE divisions
bool g_fast_rcp = true;  // refined reciprocals in the mode-batched line kernels
// the fused tail with its vectors in LDS where they fit
bool g_tail_lds = true;
bool g_cg_lds = true;  // the CG direction pass with the new directions staged in LDS
}  // namespace

// the mode-batched sweep (k_line_modes) with LP lanes per line: false when
// the level does not allow it (coupled modes, mode-dependent coefficients,
// more LDS than a workgroup gets, or a chunk length with no instance: LP of
// 32 or 64 takes rows of up to 17 per lane, 16 lanes 5..17, 8 lanes 9..33)
template <int LP, typename T>
int line_modes_rows(const MGLevT<T> &L) {
  if (L.coupled || L.qm != 0 || L.ns < 2) return 0;
  const size_t shm = (size_t)(256 / LP + 5) * lpad_len(L.ns) * sizeof(T);
  if (shm > 64 * 1024) return 0;
  const int nch = std::min(LP, L.ns / 2);
  const int rows = L.ns / nch + (L.ns % nch ? 1 : 0);
  const int lo = LP >= 32 ? 1 : (LP == 16 ? 5 : 9), hi = LP >= 32 ? 17 : (LP == 16 ? 17 : 33);
  return rows >= lo && rows <= hi ? rows : 0;
}

template <int LP, typename T>
bool launch_line_modes(hipStream_t st, const MGLevT<T> &L, const T *lam, T *x, const T *b, int col_a, int nla,
                       int zero_nb, int pin_s, const int *singular, const int *done, const T *xc, int cna) {
  const int rows = line_modes_rows<LP>(L);
  if (!rows) return false;
  constexpr int NLW = 256 / LP;
  const size_t shm = (size_t)(NLW + 5) * lpad_len(L.ns) * sizeof(T);
  dim3 gm((unsigned)((long)nla * ((L.nm + NLW - 1) / NLW)));
  auto go = [&](auto kk) {
    constexpr int KK = decltype(kk)::value;
    if (g_fast_rcp)
      hipLaunchKernelGGL((k_line_modes<KK, LP, T, true>), gm, dim3(256), shm, st, L, lam, x, b, col_a, nla, zero_nb,
                         pin_s, singular, done, xc, cna);
    else
      hipLaunchKernelGGL((k_line_modes<KK, LP, T, false>), gm, dim3(256), shm, st, L, lam, x, b, col_a, nla, zero_nb,
                         pin_s, singular, done, xc, cna);
  };
  // lines of exactly LP * K rows: the instance with compile-time chunks
  auto full = [&](auto kk) {
    constexpr int KK = decltype(kk)::value;
    if (L.ns != LP * KK) return false;
    if (g_fast_rcp)
      hipLaunchKernelGGL((k_line_modes<KK, LP, T, true, true>), gm, dim3(256), shm, st, L, lam, x, b, col_a, nla,
                         zero_nb, pin_s, singular, done, xc, cna);
    else
      hipLaunchKernelGGL((k_line_modes<KK, LP, T, false, true>), gm, dim3(256), shm, st, L, lam, x, b, col_a, nla,
                         zero_nb, pin_s, singular, done, xc, cna);
    SNAC_HIP(hipGetLastError());
    return true;
  };
  if constexpr (LP == 32) {
    if (full(std::integral_constant<int, 4>{}) || full(std::integral_constant<int, 5>{}) ||
        full(std::integral_constant<int, 8>{}))
      return true;
  } else if constexpr (LP == 16) {
    if (full(std::integral_constant<int, 8>{}) || full(std::integral_constant<int, 10>{}) ||
        full(std::integral_constant<int, 16>{}))
      return true;
  } else if constexpr (LP == 8) {
    if (full(std::integral_constant<int, 16>{}) || full(std::integral_constant<int, 20>{}) ||
        full(std::integral_constant<int, 32>{}))
      return true;
  }
  if constexpr (LP >= 32) {
    if (rows <= 3) go(std::integral_constant<int, 3>{});
    else if (rows <= 5) go(std::integral_constant<int, 5>{});
    else if (rows <= 9) go(std::integral_constant<int, 9>{});
    else go(std::integral_constant<int, 17>{});
  } else if constexpr (LP == 16) {
    if (rows <= 9) go(std::integral_constant<int, 9>{});
    else go(std::integral_constant<int, 17>{});
  } else {
    if (rows <= 17) go(std::integral_constant<int, 17>{});
    else go(std::integral_constant<int, 33>{});
  }
  SNAC_HIP(hipGetLastError());
  return true;
}

template <int LPL, typename T>
void mg_line_k(hipStream_t st, const MGLevT<T> &L, const T *lam, T *x, const T *b, int col_a, int col_m, int nla,
               int nlm, int zero_nb, int pin_s, const int *singular, const int *done, const T *xc, int cna) {
  const long lines = (long)nla * nlm;
  dim3 grid((unsigned)((lines * LPL + 255) / 256));
  const int rows = L.base + (L.extra ? 1 : 0);
  const size_t nsp = (size_t)lpad_len(L.ns);
  const size_t shb = (size_t)(256 / LPL) * nsp * sizeof(T);
  const int lv = g_line_variant >= 0 ? g_line_variant : (L.lvar >= 0 ? L.lvar : 2);
  // a variant the level does not allow falls back to the next lower one
  if (lv >= 4 && launch_line_modes<8>(st, L, lam, x, b, col_a, nla, zero_nb, pin_s, singular, done, xc, cna)) return;
  if (lv >= 3 && launch_line_modes<16>(st, L, lam, x, b, col_a, nla, zero_nb, pin_s, singular, done, xc, cna))
    return;
  if (lv >= 2 && launch_line_modes<LPL>(st, L, lam, x, b, col_a, nla, zero_nb, pin_s, singular, done, xc, cna))
    return;
  if (lv >= 1 && L.ns >= 2 && shb <= 64 * 1024) {
    if (rows <= 3)
      hipLaunchKernelGGL((k_line_lds<3, LPL, T>), grid, dim3(256), shb, st, L, lam, x, b, col_a, col_m, nla, nlm,
                         zero_nb, pin_s, singular, done, xc, cna);
    else if (rows <= 5)
      hipLaunchKernelGGL((k_line_lds<5, LPL, T>), grid, dim3(256), shb, st, L, lam, x, b, col_a, col_m, nla, nlm,
                         zero_nb, pin_s, singular, done, xc, cna);
    else if (rows <= 9)
      hipLaunchKernelGGL((k_line_lds<9, LPL, T>), grid, dim3(256), shb, st, L, lam, x, b, col_a, col_m, nla, nlm,
                         zero_nb, pin_s, singular, done, xc, cna);
    else if (rows <= 17)
      hipLaunchKernelGGL((k_line_lds<17, LPL, T>), grid, dim3(256), shb, st, L, lam, x, b, col_a, col_m, nla, nlm,
                         zero_nb, pin_s, singular, done, xc, cna);
    else
      throw std::runtime_error("multigrid line relaxation: lines longer than 1088 cells are not supported");
    SNAC_HIP(hipGetLastError());
    return;
  }
  if (rows <= 3)
    hipLaunchKernelGGL((k_line<3, LPL, T>), grid, dim3(256), 0, st, L, lam, x, b, col_a, col_m, nla, nlm, zero_nb,
                       pin_s, singular, done, xc, cna);
  else if (rows <= 5)
    hipLaunchKernelGGL((k_line<5, LPL, T>), grid, dim3(256), 0, st, L, lam, x, b, col_a, col_m, nla, nlm, zero_nb,
                       pin_s, singular, done, xc, cna);
  else if (rows <= 9)
    hipLaunchKernelGGL((k_line<9, LPL, T>), grid, dim3(256), 0, st, L, lam, x, b, col_a, col_m, nla, nlm, zero_nb,
                       pin_s, singular, done, xc, cna);
  else if (rows <= 17)
    hipLaunchKernelGGL((k_line<17, LPL, T>), grid, dim3(256), 0, st, L, lam, x, b, col_a, col_m, nla, nlm, zero_nb,
                       pin_s, singular, done, xc, cna);
  else
    throw std::runtime_error("multigrid line relaxation: lines longer than 1088 cells are not supported");
  SNAC_HIP(hipGetLastError());
}

void mg_set_line_variant(int v) { g_line_variant = v; }

template <typename T>
bool mg_line_variant_runs(const MGLevT<T> &L, int v) {
  switch (v) {
    case 0: return true;
    case 1: return L.ns >= 2 && (size_t)(256 / L.lpl) * lpad_len(L.ns) * sizeof(T) <= 64 * 1024;
    case 2: return L.lpl == 64 ? line_modes_rows<64>(L) > 0 : line_modes_rows<32>(L) > 0;
    case 3: return line_modes_rows<16>(L) > 0;
    case 4: return line_modes_rows<8>(L) > 0;
    default: return false;
  }
}
template bool mg_line_variant_runs<double>(const MGLevT<double> &, int);
template bool mg_line_variant_runs<float>(const MGLevT<float> &, int);
void mg_set_fast_rcp(bool on) { g_fast_rcp = on; }
void mg_set_tail_lds(bool on) { g_tail_lds = on; }
void mg_set_cg_lds(bool on) { g_cg_lds = on; }
bool mg_tail_lds() { return g_tail_lds; }
bool mg_fast_rcp() { return g_fast_rcp; }
int mg_line_variant() { return g_line_variant; }

template <typename T>
void mg_line(hipStream_t st, const MGLevT<T> &L, const T *lam, T *x, const T *b, int col_a, int col_m, int zero_nb,
             int pin_s, const int *singular, const int *done, const T *xc, int cna) {
  const int nla = (L.na - col_a + 1) / 2;
  const int nlm = L.coupled ? (L.nm - col_m + 1) / 2 : L.nm;
  if (nla <= 0 || nlm <= 0) return;
  if (xc && (L.coupled || zero_nb)) throw std::runtime_error("internal: fused prolongation needs an FFT-path sweep");
  if (L.lpl == 32) mg_line_k<32>(st, L, lam, x, b, col_a, col_m, nla, nlm, zero_nb, pin_s, singular, done, xc, cna);
  else mg_line_k<64>(st, L, lam, x, b, col_a, col_m, nla, nlm, zero_nb, pin_s, singular, done, xc, cna);
}

template <int LPL, typename T>
void mg_tail_k(hipStream_t st, const MGTailT<T> &t, int nm, const T *lam, int pin_s, const int *singular,
               const int *done) {
  const MGLevT<T> &L = t.lev[0];
  const int rows = L.base + (L.extra ? 1 : 0);
  const size_t shm = (size_t)t.lds_elems * sizeof(T);
  const bool lx = g_tail_lds && shm <= 160 * 1024;
  // the fused coarse end divides by IEEE quotients whatever the line
  // kernels do (its compile-time-chunk instances are the fastest form)
#define LAUNCH_TAIL(KK)                                                                                             \
  do {                                                                                                              \
    if (lx)                                                                                                         \
      hipLaunchKernelGGL((k_tail<KK, LPL, T, false, true>), dim3(nm), dim3(1024), shm, st, t, lam, pin_s, singular,  \
                         done);                                                                                     \
    else                                                                                                            \
      hipLaunchKernelGGL((k_tail<KK, LPL, T, false, false>), dim3(nm), dim3(1024), 0, st, t, lam, pin_s, singular,   \
                         done);                                                                                     \
  } while (0)
  // lines of exactly LPL * K rows (IEEE divisions): the instance with
  // compile-time chunks
  bool done_full = false;
  if constexpr (LPL == 32) {
    auto full = [&](auto kk) {
      constexpr int KK = decltype(kk)::value;
      if (L.ns != LPL * KK) return false;
      if (lx)
        hipLaunchKernelGGL((k_tail<KK, LPL, T, false, true, true>), dim3(nm), dim3(1024), shm, st, t, lam, pin_s,
                           singular, done);
      else
        hipLaunchKernelGGL((k_tail<KK, LPL, T, false, false, true>), dim3(nm), dim3(1024), 0, st, t, lam, pin_s,
                           singular, done);
      return true;
    };
    done_full = full(std::integral_constant<int, 4>{}) || full(std::integral_constant<int, 5>{}) ||
                full(std::integral_constant<int, 8>{});
  }
  if (done_full) {
  } else if (rows <= 3) LAUNCH_TAIL(3);
  else if (rows <= 5) LAUNCH_TAIL(5);
  else if (rows <= 9) LAUNCH_TAIL(9);
  else LAUNCH_TAIL(17);
#undef LAUNCH_TAIL
  SNAC_HIP(hipGetLastError());
}

template <typename T>
void mg_tail(hipStream_t st, const MGTailT<T> &t, int nm, const T *lam, int pin_s, const int *singular,
             const int *done) {
  if (t.lev[0].lpl == 32) mg_tail_k<32>(st, t, nm, lam, pin_s, singular, done);
  else mg_tail_k<64>(st, t, nm, lam, pin_s, singular, done);
}

template <typename T>
void mg_restrict_residual(hipStream_t st, const MGLevT<T> &F, const MGLevT<T> &C, const T *lam_f, const T *x,
                          const T *b, T *bc, const int *done) {
  const long n = (long)C.nm * C.na * C.ns;
  constexpr int RM = 8;
  const size_t shm = (size_t)2 * F.ns * (5 * sizeof(T) + 1);
  const int lv = g_line_variant >= 0 ? g_line_variant : (F.lvar >= 0 ? F.lvar : 2);
  if (lv >= 2 && !F.coupled && F.qm == 0 && F.nm == C.nm && shm <= 64 * 1024) {
    dim3 g((unsigned)((long)C.na * ((C.nm + RM - 1) / RM)));
    hipLaunchKernelGGL((k_restrict_modes<T, RM>), g, dim3(256), shm, st, F, C, lam_f, x, b, bc, done);
    SNAC_HIP(hipGetLastError());
    return;
  }
  hipLaunchKernelGGL(k_restrict<T>, dim3(blocks_for(n)), dim3(256), 0, st, F, C, lam_f, x, b, bc, done);
  SNAC_HIP(hipGetLastError());
}

template <typename T>
void mg_prolong(hipStream_t st, const MGLevT<T> &F, const MGLevT<T> &C, const T *xc, T *xf, const int *done) {
  const long n = (long)F.nm * F.na * F.ns;
  hipLaunchKernelGGL(k_prolong<T>, dim3(blocks_for(n)), dim3(256), 0, st, F, C, xc, xf, done);
  SNAC_HIP(hipGetLastError());
}

template <typename T>
void mg_zero(hipStream_t st, long n, T *x) {
  hipLaunchKernelGGL(k_zero<T>, dim3(blocks_for(n)), dim3(256), 0, st, n, x);
  SNAC_HIP(hipGetLastError());
}

template <typename T>
void mg_pattern(hipStream_t st, long n, T *x) {
  hipLaunchKernelGGL(k_pattern<T>, dim3(blocks_for(n)), dim3(256), 0, st, n, x);
  SNAC_HIP(hipGetLastError());
}

void mg_apply_dot(hipStream_t st, const MGLev &L, const double *lam, const double *p, double *q, int ng, long gsize,
                  double *part, int nchunk, double *out, const int *done) {
  hipLaunchKernelGGL(k_apply_dot, dim3(nchunk, ng), dim3(256), 0, st, L, lam, p, q, part, nchunk, gsize, done);
  SNAC_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_dot_final, dim3(ng), dim3(64), 0, st, part, nchunk, out, done, nullptr);
  SNAC_HIP(hipGetLastError());
}

void mg_warm_residual(hipStream_t st, int ng, long gsize, const double *bb, double *x, double *r, const double *q,
                      float *rf, double *part, int nchunk, double *out) {
  hipLaunchKernelGGL(k_warm_residual, dim3(nchunk, ng), dim3(256), 0, st, gsize, bb, x, r, q, rf, part, nchunk);
  SNAC_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_dot_final, dim3(ng), dim3(64), 0, st, part, nchunk, out, nullptr, nullptr);
  SNAC_HIP(hipGetLastError());
}

void mg_cg_update_dot(hipStream_t st, int ng, long gsize, const double *rz, const double *pq, double *x,
                      const double *p, double *r, const double *q, float *rf, const float *zf, double *part,
                      int nchunk, double *out, const int *done) {
  hipLaunchKernelGGL(k_cg_update_dot, dim3(nchunk, ng), dim3(256), 0, st, gsize, rz, pq, x, p, r, q, rf, zf, part,
                     nchunk, done);
  SNAC_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_dot_final, dim3(ng), dim3(64), 0, st, part, nchunk, out, done, nullptr);
  SNAC_HIP(hipGetLastError());
}

template <typename TZ, typename TP>
void mg_dir_apply_dot(hipStream_t st, const MGLev &L, const double *lam, int ng, long gsize, const double *rzn,
                      const double *rzo, const double *cross, const TZ *z, const TP *p_prev, TP *p, double *part,
                      int nchunk, double *out, const int *done) {
  // the LDS form where a level's coefficients do not depend on a, its rows
  // divide 256 and a chunk's lines with their two neighbours fit in 64 KB
  const long csz = ((gsize + nchunk - 1) / nchunk + 255) / 256 * 256;
  const size_t lds = L.ns > 0 ? (size_t)(csz / L.ns + 2) * L.ns * sizeof(TP) : 0;
  const size_t ldsg = (size_t)(csz + 2 * L.ns) * sizeof(TP);  // any 2D level: the chunk and a line on each side
  if (g_cg_lds && !L.coupled && L.auni && L.ns > 0 && 256 % L.ns == 0 && lds <= 64 * 1024)
    hipLaunchKernelGGL((k_dir_apply_dot_lds<TZ, TP>), dim3(nchunk, ng), dim3(256), lds, st, L, lam, gsize, rzn, rzo,
                       cross, z, p_prev, p, part, nchunk, done);
  else if (g_cg_lds && !L.coupled && ldsg <= 64 * 1024)
    hipLaunchKernelGGL((k_dir_apply_dot<TZ, TP, true>), dim3(nchunk, ng), dim3(256), ldsg, st, L, lam, gsize, rzn, rzo,
                       cross, z, p_prev, p, part, nchunk, done);
  else
    hipLaunchKernelGGL((k_dir_apply_dot<TZ, TP>), dim3(nchunk, ng), dim3(256), 0, st, L, lam, gsize, rzn, rzo, cross,
                       z, p_prev, p, part, nchunk, done);
  SNAC_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_dot_final, dim3(ng), dim3(64), 0, st, part, nchunk, out, done, nullptr);
  SNAC_HIP(hipGetLastError());
}

template <typename TP>
void mg_update_apply_dot(hipStream_t st, const MGLev &L, const double *lam, int ng, long gsize, const double *rz,
                         const double *pq, double *x, const TP *p, double *r, float *rf, const float *zf,
                         double *part, int nchunk, double *out, const int *done, bool first) {
  const long csz = ((gsize + nchunk - 1) / nchunk + 255) / 256 * 256;
  const size_t lds = L.ns > 0 ? (size_t)(csz / L.ns + 2) * L.ns * sizeof(TP) : 0;
  const size_t ldsg = (size_t)(csz + 2 * L.ns) * sizeof(TP);
  const int fi = first ? 1 : 0;
  if (g_cg_lds && !L.coupled && L.auni && L.ns > 0 && 256 % L.ns == 0 && lds <= 64 * 1024)
    hipLaunchKernelGGL(k_update_apply_dot_lds<TP>, dim3(nchunk, ng), dim3(256), lds, st, L, lam, gsize, rz, pq, x, p,
                       r, rf, zf, part, nchunk, done, fi);
  else if (g_cg_lds && !L.coupled && ldsg <= 64 * 1024)
    hipLaunchKernelGGL((k_update_apply_dot<TP, true>), dim3(nchunk, ng), dim3(256), ldsg, st, L, lam, gsize, rz, pq, x,
                       p, r, rf, zf, part, nchunk, done, fi);
  else
    hipLaunchKernelGGL(k_update_apply_dot<TP>, dim3(nchunk, ng), dim3(256), 0, st, L, lam, gsize, rz, pq, x, p, r, rf,
                       zf, part, nchunk, done, fi);
  SNAC_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_dot_final, dim3(ng), dim3(64), 0, st, part, nchunk, out, done, nullptr);
  SNAC_HIP(hipGetLastError());
}
template void mg_dir_apply_dot<double, double>(hipStream_t, const MGLev &, const double *, int, long, const double *,
                                               const double *, const double *, const double *, const double *,
                                               double *, double *, int, double *, const int *);
template void mg_dir_apply_dot<float, double>(hipStream_t, const MGLev &, const double *, int, long,
                                              const double *, const double *, const double *, const float *,
                                              const double *, double *, double *, int, double *, const int *);
template void mg_dir_apply_dot<float, float>(hipStream_t, const MGLev &, const double *, int, long, const double *,
                                             const double *, const double *, const float *, const float *, float *,
                                             double *, int, double *, const int *);
template void mg_update_apply_dot<double>(hipStream_t, const MGLev &, const double *, int, long, const double *,
                                          const double *, double *, const double *, double *, float *, const float *,
                                          double *, int, double *, const int *, bool);
template void mg_update_apply_dot<float>(hipStream_t, const MGLev &, const double *, int, long, const double *,
                                         const double *, double *, const float *, double *, float *, const float *,
                                         double *, int, double *, const int *, bool);

void mg_set_shift(MGPlan &P, hipStream_t st, double shift) {
  if (P.comp < 0 || !P.lam_base) return;
  // the zebra sweeps alone contract the error by about (cross / (|shift| w))^2
  P.smooth_only = P.cross_max <= 0.1 * std::fabs(shift);
  hipLaunchKernelGGL(k_shift_lam, dim3((P.ng + 255) / 256), dim3(256), 0, st, P.ng, P.lam_base, shift, P.lam64,
                     P.lam32);
  SNAC_HIP(hipGetLastError());
}

// r.r per group (k_dot's element order and sums, so the same values) with
// rf = (float) r written in the same pass
__global__ __launch_bounds__(256) void k_dot_self_cvt(long gsize, const double *__restrict__ x, float *__restrict__ xf,
                                                      double *__restrict__ part, int nchunk) {
  __shared__ double sh[2][4];
  const int g = blockIdx.y;
  double s1 = 0.0;
  const long base = (long)g * gsize;
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < gsize; i += (long)nchunk * 256) {
    const double v = x[base + i];
    s1 += v * v;
    xf[base + i] = (float)v;
  }
  s1 = wave_sum(s1);
  if ((threadIdx.x & 63) == 0) sh[0][threadIdx.x >> 6] = s1;
  __syncthreads();
  if (threadIdx.x == 0) {
    part[((long)g * nchunk + blockIdx.x) * 2] = sh[0][0] + sh[0][1] + sh[0][2] + sh[0][3];
    part[((long)g * nchunk + blockIdx.x) * 2 + 1] = 0.0;
  }
}

void mg_dot_self_cvt(hipStream_t st, int ng, long gsize, const double *x, float *xf, double *part, int nchunk,
                     double *out) {
  hipLaunchKernelGGL(k_dot_self_cvt, dim3(nchunk, ng), dim3(256), 0, st, gsize, x, xf, part, nchunk);
  SNAC_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_dot_final, dim3(ng), dim3(64), 0, st, part, nchunk, out, nullptr, nullptr);
  SNAC_HIP(hipGetLastError());
}

void mg_to_float(hipStream_t st, long n, const double *src, float *dst) {
  hipLaunchKernelGGL(k_to_float, dim3(blocks_for(n)), dim3(256), 0, st, n, src, dst);
  SNAC_HIP(hipGetLastError());
}

void mg_izero(hipStream_t st, long n, int *x) {
  hipLaunchKernelGGL(k_izero, dim3(blocks_for(n)), dim3(256), 0, st, n, x);
  SNAC_HIP(hipGetLastError());
}

template <typename TY, typename TX2>
void mg_dot(hipStream_t st, int ng, long gsize, const double *x, const TY *y, const TX2 *x2, const double *y2,
            double *part, int nchunk, double *out, const int *done, const int *only) {
  hipLaunchKernelGGL((k_dot<TY, TX2>), dim3(nchunk, ng), dim3(256), 0, st, gsize, x, y, x2, y2, part, nchunk, done, only);
  SNAC_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_dot_final, dim3(ng), dim3(64), 0, st, part, nchunk, out, done, only);
  SNAC_HIP(hipGetLastError());
}

template <typename T>
void mg_cg_dir(hipStream_t st, int ng, long gsize, const double *rzn, const double *rzo, const double *cross,
               const T *z, double *p, const int *done) {
  hipLaunchKernelGGL(k_cg_dir<T>, dim3(blocks_for((long)ng * gsize)), dim3(256), 0, st, ng, gsize, rzn, rzo, cross, z,
                     p, done);
  SNAC_HIP(hipGetLastError());
}

template <typename T>
void mg_sub_mean(hipStream_t st, const MGLev &L, T *x, const double *sums, const int *glist, int nlist, long gsize,
                 const double *nact, const int *done) {
  if (nlist <= 0) return;
  hipLaunchKernelGGL(k_sub_mean<T>, dim3((unsigned)((gsize + 255) / 256), nlist), dim3(256), 0, st, L, x, sums, glist,
                     gsize, nact, done);
  SNAC_HIP(hipGetLastError());
}

void mg_flags(hipStream_t st, int ng, const double *rr, const double *bb, double tol2, int *done, int *left,
              int *list, int it) {
  hipLaunchKernelGGL(k_flags, dim3(1), dim3(256), 0, st, ng, rr, bb, tol2, done, left, list, it);
  SNAC_HIP(hipGetLastError());
}

void mg_copy3(hipStream_t st, const double *src, long sm, long sa, long ss, double *dst, long dm, long da, long ds,
              int em, int ea, int es, double scale, int tile) {
  if (em <= 0 || ea <= 0 || es <= 0) return;
  if (tile == 0) {
    hipLaunchKernelGGL(k_copy3<0>, dim3((es + 255) / 256, ea, em), dim3(256), 0, st, src, sm, sa, ss, dst, dm, da, ds,
                       ea, es, scale);
  } else {
    dim3 grid((es + TT - 1) / TT, (ea + TT - 1) / TT, em);
    if (tile == 1)
      hipLaunchKernelGGL(k_copy3<1>, grid, dim3(256), 0, st, src, sm, sa, ss, dst, dm, da, ds, ea, es, scale);
    else
      hipLaunchKernelGGL(k_copy3<2>, grid, dim3(256), 0, st, src, sm, sa, ss, dst, dm, da, ds, ea, es, scale);
  }
  SNAC_HIP(hipGetLastError());
}

#define SNAC_MG_INST(T)                                                                                               \
  template void mg_line<T>(hipStream_t, const MGLevT<T> &, const T *, T *, const T *, int, int, int, int, const int *, \
                           const int *, const T *, int);                                                              \
  template void mg_tail<T>(hipStream_t, const MGTailT<T> &, int, const T *, int, const int *, const int *);           \
  template void mg_restrict_residual<T>(hipStream_t, const MGLevT<T> &, const MGLevT<T> &, const T *, const T *,       \
                                        const T *, T *, const int *);                                                 \
  template void mg_prolong<T>(hipStream_t, const MGLevT<T> &, const MGLevT<T> &, const T *, T *, const int *);        \
  template void mg_zero<T>(hipStream_t, long, T *);                                                                   \
  template void mg_pattern<T>(hipStream_t, long, T *);                                                                   \
  template void mg_dot<T, T>(hipStream_t, int, long, const double *, const T *, const T *, const double *, double *,  \
                             int, double *, const int *, const int *);                                                \
  template void mg_cg_dir<T>(hipStream_t, int, long, const double *, const double *, const double *, const T *,       \
                             double *, const int *);                                                                  \
  template void mg_sub_mean<T>(hipStream_t, const MGLev &, T *, const double *, const int *, int, long,             \
                               const double *, const int *);
SNAC_MG_INST(double)
SNAC_MG_INST(float)
#undef SNAC_MG_INST
template void mg_dot<double, float>(hipStream_t, int, long, const double *, const double *, const float *,
                                    const double *, double *, int, double *, const int *, const int *);

}  // namespace snac
