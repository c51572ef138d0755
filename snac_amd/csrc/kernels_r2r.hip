// Batched real-to-real transforms along one axis, the kinds FFTW offers and
// SNaC selects by boundary condition (fft.f90:96-149): R2HC/HC2R (periodic),
// REDFT10/REDFT01 (Neumann-Neumann), RODFT10/RODFT01 (Dirichlet-Dirichlet);
// REDFT11/RODFT11 (mixed) and lengths that are not powers of two go through
// a dense matrix path.  Unnormalised, as FFTW's; callers scale.
//
// FFT path (n = 4..1024, powers of two): a workgroup stages L whole lines in
// LDS, forms the kind's n-point complex sequence (halfcomplex unpacking for
// HC2R, Makhoul's even/odd reordering for the cosine and sine kinds), and
// runs a two-pass Cooley-Tukey split n = N1*N2: every thread does one N1- or
// N2-point DFT entirely in registers, with one LDS transpose in between
// instead of log2(n) radix-2 stages through LDS.  Loads and stores walk the
// batch index with unit stride when the line stride is not 1, so a row of L
// neighbouring lines is one coalesced segment.
#include <cmath>
#include <map>
#include <mutex>
#include <vector>

#include "common.hpp"

namespace snac {
namespace {

constexpr int NT = 256;
constexpr int LDS_COMPLEX = 4096;  // complex doubles per workgroup (64 KiB)

enum { K_R2HC = 0, K_HC2R = 1, K_REDFT00 = 3, K_REDFT01 = 4, K_REDFT10 = 5, K_REDFT11 = 6, K_RODFT00 = 7,
       K_RODFT01 = 8, K_RODFT10 = 9, K_RODFT11 = 10 };

struct Args {
  const double *in;
  double *out;
  long ies, is1, is2, oes, os1, os2;
  int n, logn, nb1, nlines, L;
  int coal;  // 1: consecutive threads walk the batch index (line stride != 1)
  double scale;
  const double *tab;
};

// lines per workgroup for an N1 x N2 FFT: one task per thread in each pass
// and at most LDS_COMPLEX complex values of LDS
constexpr int fft_lines(int n1, int n2) {
  return (NT / (n1 > n2 ? n1 : n2)) < (LDS_COMPLEX / (n1 * n2)) ? (NT / (n1 > n2 ? n1 : n2))
                                                                : (LDS_COMPLEX / (n1 * n2));
}

// Input / output offsets of the workgroup's L lines, one division each,
// kept in LDS so the per-element address is one multiply-add (-1: past the
// last line).  Ends with a barrier.
template <int L>
__device__ __forceinline__ void line_offsets(const Args &a, int line0, long (&oin)[L], long (&oout)[L]) {
  for (int t = threadIdx.x; t < L; t += NT) {
    const int line = line0 + t;
    long oi = -1, oo = -1;
    if (line < a.nlines) {
      const int b1 = line % a.nb1, b2 = line / a.nb1;
      oi = b1 * a.is1 + b2 * a.is2;
      oo = b1 * a.os1 + b2 * a.os2;
    }
    oin[t] = oi;
    oout[t] = oo;
  }
  __syncthreads();
}

__device__ __forceinline__ void elem_of(const Args &a, int e, int &line, int &q) {
  if (a.coal) { line = e % a.L; q = e / a.L; }
  else { q = e % a.n; line = e / a.n; }
}

// radix-2 DIT on N values held in registers (fully unrolled; twiddle
// indices are compile-time, so the loads are uniform across the wavefront).
// tw[m] = e^{-2 pi i m/n}; stride = n/N maps W_N^m onto that table.
template <int N>
__device__ __forceinline__ void fft_reg(double2 (&v)[N], const double2 *__restrict__ tw, int stride, bool inv) {
  constexpr int LOG = (N >= 2) + (N >= 4) + (N >= 8) + (N >= 16) + (N >= 32);
#pragma unroll
  for (int i = 0; i < N; i++) {
    int j = 0;
#pragma unroll
    for (int b = 0; b < LOG; b++) j |= ((i >> b) & 1) << (LOG - 1 - b);
    if (j > i) { double2 t = v[i]; v[i] = v[j]; v[j] = t; }
  }
#pragma unroll
  for (int len = 2; len <= N; len <<= 1) {
#pragma unroll
    for (int i = 0; i < N; i += len) {
#pragma unroll
      for (int j = 0; j < len / 2; j++) {
        double2 w = tw[j * (N / len) * stride];
        if (inv) w.y = -w.y;
        double2 x0 = v[i + j], x1 = v[i + j + len / 2];
        double2 c = make_double2(x1.x * w.x - x1.y * w.y, x1.x * w.y + x1.y * w.x);
        v[i + j] = make_double2(x0.x + c.x, x0.y + c.y);
        v[i + j + len / 2] = make_double2(x0.x - c.x, x0.y - c.y);
      }
    }
  }
}

// complex input slot s of the n-point sequence for KIND, from the raw line r
// (raw index n maps to zero where the kind needs x_n).  tk[k] = (cos, sin)(pi k / 2n).
template <int KIND>
__device__ __forceinline__ double2 pre_slot(const double *r, int s, int n, const double2 *tk) {
  const int h = n >> 1;
  if (KIND == K_R2HC) return make_double2(r[s], 0.0);
  if (KIND == K_HC2R) {
    if (s == 0 || s == h) return make_double2(r[s], 0.0);
    if (s < h) return make_double2(r[s], r[n - s]);
    return make_double2(r[n - s], -r[s]);
  }
  if (KIND == K_REDFT10 || KIND == K_RODFT10) {
    int q = (s < h) ? 2 * s : 2 * (n - 1 - s) + 1;
    double x = r[q];
    if (KIND == K_RODFT10 && (q & 1)) x = -x;
    return make_double2(x, 0.0);
  }
  // REDFT01 / RODFT01: e^{i pi s/(2n)} (a - i b)
  double a, b;
  if (KIND == K_REDFT01) { a = r[s]; b = (s == 0) ? 0.0 : r[n - s]; }
  else { a = r[n - 1 - s]; b = (s == 0) ? 0.0 : r[s - 1]; }
  double2 t = tk[s];
  return make_double2(a * t.x + b * t.y, a * t.y - b * t.x);
}

template <int KIND>
__device__ __forceinline__ double post_val(const double2 *z, int q, int n, const double2 *tk) {
  if (KIND == K_R2HC) return (q <= (n >> 1)) ? z[q].x : z[n - q].y;
  if (KIND == K_HC2R) return z[q].x;
  if (KIND == K_REDFT10 || KIND == K_RODFT10) {
    int k = (KIND == K_RODFT10) ? n - 1 - q : q;
    double2 t = tk[k], v = z[k];
    return 2.0 * (v.x * t.x + v.y * t.y);
  }
  int slot = (q & 1) ? n - 1 - (q >> 1) : (q >> 1);
  double y = z[slot].x;
  if (KIND == K_RODFT01 && (q & 1)) y = -y;
  return y;
}

// n = N1*N2 points per line; j = N2*j1 + j2, k = k1 + N1*k2:
//   Y[j2][k1] = sum_j1 z[N2 j1 + j2] W_N1^{j1 k1}, times W_n^{j2 k1}
//   Z[k1 + N1 k2] = sum_j2 Y[j2][k1] W_N2^{j2 k2}
// One task per thread in each pass (the host sizes L so that L*N1 and L*N2
// do not exceed the workgroup).  Y is stored row k1 with its N2 entries
// XOR-swizzled by k1, so both the pass-1 writes (consecutive j2) and the
// pass-2 reads (consecutive k1) spread over the LDS banks.  All IT loads of
// a thread are issued before its first LDS write.
template <int N1, int N2, int KIND, int LSHIFT>
__global__ __launch_bounds__(NT) void k_r2r_fft(Args a) {
  extern __shared__ double2 S[];
  constexpr int n = N1 * N2;
  constexpr int L = (fft_lines(N1, N2) >> LSHIFT) > 0 ? (fft_lines(N1, N2) >> LSHIFT) : 1;
  constexpr int IT = (L * n + NT - 1) / NT;  // elements per thread in the load / store phases
  const int line0 = blockIdx.x * L;
  const double2 *tw = reinterpret_cast<const double2 *>(a.tab);      // e^{-2 pi i m/n}, m < n
  const double2 *tk = reinterpret_cast<const double2 *>(a.tab) + n;  // (cos, sin)(pi k/2n)
  constexpr bool inv = (KIND == K_HC2R || KIND == K_REDFT01 || KIND == K_RODFT01);
  double *R = reinterpret_cast<double *>(S);
  __shared__ long oin[L], oout[L];
  line_offsets<L>(a, line0, oin, oout);
  // 1. raw lines into LDS, line l at doubles [2 n l, 2 n l + n)
  {
    double xr[IT];
#pragma unroll
    for (int it = 0; it < IT; it++) {
      const int e = threadIdx.x + it * NT;
      const int l = a.coal ? e % L : e / n, q = a.coal ? e / L : e % n;
      xr[it] = 0.0;
      if (e < L * n && oin[l] >= 0) xr[it] = a.in[q * a.ies + oin[l]];
    }
#pragma unroll
    for (int it = 0; it < IT; it++) {
      const int e = threadIdx.x + it * NT;
      const int l = a.coal ? e % L : e / n, q = a.coal ? e / L : e % n;
      if (e < L * n) R[2 * n * l + q] = xr[it];
    }
  }
  __syncthreads();
  // 2. first pass: N1-point DFTs over j1 for every (line, j2)
  {
    const int t = threadIdx.x;
    const bool act = t < L * N2;
    const int j2 = t % N2, l = t / N2;
    double2 v[N1];
    if (act) {
      const double *r = R + 2 * n * l;
#pragma unroll
      for (int j1 = 0; j1 < N1; j1++) v[j1] = pre_slot<KIND>(r, N2 * j1 + j2, n, tk);
      fft_reg<N1>(v, tw, N2, inv);
#pragma unroll
      for (int k1 = 1; k1 < N1; k1++) {
        double2 w = tw[(j2 * k1) & (n - 1)];
        if (inv) w.y = -w.y;
        double2 y = v[k1];
        v[k1] = make_double2(y.x * w.x - y.y * w.y, y.x * w.y + y.y * w.x);
      }
    }
    __syncthreads();
    if (act) {
      double2 *T = S + n * l;
#pragma unroll
      for (int k1 = 0; k1 < N1; k1++) T[k1 * N2 + (j2 ^ (k1 & (N2 - 1)))] = v[k1];
    }
    __syncthreads();
  }
  // 3. second pass: N2-point DFTs over j2 for every (line, k1)
  {
    const int t = threadIdx.x;
    const bool act = t < L * N1;
    const int k1 = t % N1, l = t / N1;
    double2 v[N2];
    if (act) {
      const double2 *T = S + n * l;
#pragma unroll
      for (int j2 = 0; j2 < N2; j2++) v[j2] = T[k1 * N2 + (j2 ^ (k1 & (N2 - 1)))];
      fft_reg<N2>(v, tw, N1, inv);
    }
    __syncthreads();
    if (act) {
      double2 *Z = S + n * l;
#pragma unroll
      for (int k2 = 0; k2 < N2; k2++) Z[k1 + N1 * k2] = v[k2];
    }
    __syncthreads();
  }
  // 4. post-mapping and store
#pragma unroll
  for (int it = 0; it < IT; it++) {
    const int e = threadIdx.x + it * NT;
    const int l = a.coal ? e % L : e / n, q = a.coal ? e / L : e % n;
    if (e < L * n && oout[l] >= 0) {
      const double y = post_val<KIND>(S + n * l, q, n, tk);
      a.out[q * a.oes + oout[l]] = y * a.scale;
    }
  }
}

// R2HC / HC2R of n = 2h points through one h-point complex FFT per line
// (h = N1*N2), half the work and half the LDS of the n-point form:
//   forward: z[m] = x[2m] + i x[2m+1]; Z = FFT_h(z); for k = 0..h
//     X[k] = E + W_n^k O,  E = (Z[k] + conj Z[h-k]) / 2,  O = (Z[k] - conj Z[h-k]) / 2i
//   inverse: Z[k] = A + i B,  A = X[k] + conj X[h-k],  B = (X[k] - conj X[h-k]) w^k
//     (w = e^{2 pi i/n}), z = IFFT_h(Z), y[2m] + i y[2m+1] = z[m]
// (indices of Z modulo h).  Every line is transformed on its own, so a line
// of zeros stays exactly zero.  The raw line and its h complex values share
// one n-double slot of LDS.
#ifndef SNAC_RFFT_PAD
#define SNAC_RFFT_PAD 2
#endif
constexpr int RPAD = SNAC_RFFT_PAD;  // doubles of padding after each LDS line (spreads lines over the banks)
constexpr int rfft_lines(int n1, int n2) {
  return (NT / (n1 > n2 ? n1 : n2)) < (LDS_COMPLEX / (2 * n1 * n2)) ? (NT / (n1 > n2 ? n1 : n2))
                                                                    : (LDS_COMPLEX / (2 * n1 * n2));
}

// The two passes of the h-point complex FFT of one line (h = N1*N2) and
// its unfolding.  The line's LDS slot holds the n = 2h raw doubles (r) and,
// after pass 1, its h complex values (T, XOR-swizzled between the passes,
// then Z in natural order).  k_rfft_w uses these; k_rfft keeps the same
// steps written inline, which the compiler allocates in fewer registers
// there (164 against 172 VGPRs for the n = 256 inverse: a wave per SIMD).
// Pass 1: N1-point DFT over j1 of column j2, the inverse's folding of the
// halfcomplex input on the way in, the twiddles W_h^{j2 k1} on the way out.
template <int N1, int N2, bool INV>
__device__ __forceinline__ void rfft_pass1(const double *r, int j2, const double2 *tw, double2 (&v)[N1]) {
  constexpr int h = N1 * N2, n = 2 * h;
#pragma unroll
  for (int j1 = 0; j1 < N1; j1++) {
    const int k = N2 * j1 + j2;
    if (!INV) {
      v[j1] = make_double2(r[2 * k], r[2 * k + 1]);
    } else {
      // X[k] and conj X[h-k] from the halfcomplex line
      const double2 xk = (k == 0) ? make_double2(r[0], 0.0) : make_double2(r[k], r[n - k]);
      const double2 xc = (k == 0) ? make_double2(r[h], 0.0) : make_double2(r[h - k], -r[h + k]);
      const double2 w = tw[k];  // conj of w^k
      const double dx = xk.x - xc.x, dy = xk.y - xc.y;
      const double bx = dx * w.x + dy * w.y, by = dy * w.x - dx * w.y;  // (dx + i dy) e^{+2 pi i k/n}
      v[j1] = make_double2(xk.x + xc.x - by, xk.y + xc.y + bx);        // A + i B
    }
  }
  fft_reg<N1>(v, tw, 2 * N2, INV);
#pragma unroll
  for (int k1 = 1; k1 < N1; k1++) {
    double2 w = tw[(2 * j2 * k1) & (n - 1)];
    if (INV) w.y = -w.y;
    const double2 y = v[k1];
    v[k1] = make_double2(y.x * w.x - y.y * w.y, y.x * w.y + y.y * w.x);
  }
}
template <int N1, int N2>
__device__ __forceinline__ void rfft_put1(double2 *T, int j2, const double2 (&v)[N1]) {
#pragma unroll
  for (int k1 = 0; k1 < N1; k1++) T[k1 * N2 + (j2 ^ (k1 & (N2 - 1)))] = v[k1];
}
// Pass 2: N2-point DFT over j2 of row k1
template <int N1, int N2, bool INV>
__device__ __forceinline__ void rfft_pass2(const double2 *T, int k1, const double2 *tw, double2 (&v)[N2]) {
#pragma unroll
  for (int j2 = 0; j2 < N2; j2++) v[j2] = T[k1 * N2 + (j2 ^ (k1 & (N2 - 1)))];
  fft_reg<N2>(v, tw, 2 * N1, INV);
}
template <int N1, int N2>
__device__ __forceinline__ void rfft_put2(double2 *Z, int k1, const double2 (&v)[N2]) {
#pragma unroll
  for (int k2 = 0; k2 < N2; k2++) Z[k1 + N1 * k2] = v[k2];
}
// Output q of the line from its h-point spectrum Z: forward, the halfcomplex
// X[q] (q <= h: real part, else the imaginary part of X[n-q]); inverse, the
// real sample y[q]
template <int N1, int N2, bool INV>
__device__ __forceinline__ double rfft_out(const double2 *Z, int q, const double2 *tw) {
  constexpr int h = N1 * N2, n = 2 * h;
  if (!INV) {
    const int k = q <= h ? q : n - q;
    const double2 zk = Z[k & (h - 1)], zc = Z[(h - k) & (h - 1)];  // Z[k], Z[h-k] (mod h)
    const double ex = 0.5 * (zk.x + zc.x), ey = 0.5 * (zk.y - zc.y);
    const double ox = 0.5 * (zk.y + zc.y), oy = -0.5 * (zk.x - zc.x);
    const double2 w = tw[k];
    const double xr = ex + (w.x * ox - w.y * oy), xi = ey + (w.x * oy + w.y * ox);
    return q <= h ? xr : xi;
  }
  const double2 z = Z[q >> 1];
  return (q & 1) ? z.y : z.x;
}

// TPB tiles of L lines per workgroup: the loads of tile t+1 are issued
// before tile t is transformed and land in LDS after it, so their latency
// overlaps the arithmetic (TPB = 1: one tile, no overlap).
template <int N1, int N2, bool INV, int LSHIFT, int TPB>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(N1 < N2 ? 4 : 1))) void k_rfft(Args a) {
  extern __shared__ double2 S[];
  constexpr int h = N1 * N2, n = 2 * h;
  constexpr int L = (rfft_lines(N1, N2) >> LSHIFT) > 0 ? (rfft_lines(N1, N2) >> LSHIFT) : 1;
  constexpr int IT = (L * n + NT - 1) / NT;
  const double2 *tw = reinterpret_cast<const double2 *>(a.tab);  // e^{-2 pi i m/n}, m < n
  constexpr int LS = n + RPAD;                                     // LDS doubles per line
  double *R = reinterpret_cast<double *>(S);                     // line l at doubles [LS l, LS l + n)
  __shared__ long oin[2][L], oout[2][L];
  const int ntile = (a.nlines + L - 1) / L, tile0 = blockIdx.x * TPB;
  auto offsets = [&](int slot, int tile) {
    for (int t = threadIdx.x; t < L; t += NT) {
      const int line = tile * L + t;
      long oi = -1, oo = -1;
      if (line < a.nlines) {
        const int b1 = line % a.nb1, b2 = line / a.nb1;
        oi = b1 * a.is1 + b2 * a.is2;
        oo = b1 * a.os1 + b2 * a.os2;
      }
      oin[slot][t] = oi;
      oout[slot][t] = oo;
    }
  };
  double xr[IT];
  auto load = [&](int slot) {
#pragma unroll
    for (int it = 0; it < IT; it++) {
      const int e = threadIdx.x + it * NT;
      const int l = a.coal ? e % L : e / n, q = a.coal ? e / L : e % n;
      xr[it] = 0.0;
      if (e < L * n && oin[slot][l] >= 0) xr[it] = a.in[q * a.ies + oin[slot][l]];
    }
  };
  offsets(0, tile0);
  __syncthreads();
  load(0);
  for (int tt = 0; tt < TPB; tt++) {
  const int tile = tile0 + tt;
  if (tile >= ntile) break;  // uniform over the workgroup
  const int sl = tt & 1;
  const bool more = tt + 1 < TPB && tile + 1 < ntile;
  __syncthreads();  // the previous tile is done with the LDS lines and its offsets
  // 1. raw lines into LDS
#pragma unroll
  for (int it = 0; it < IT; it++) {
    const int e = threadIdx.x + it * NT;
    const int l = a.coal ? e % L : e / n, q = a.coal ? e / L : e % n;
    if (e < L * n) R[LS * l + q] = xr[it];
  }
  if (more) offsets(sl ^ 1, tile + 1);
  __syncthreads();
  if (more) load(sl ^ 1);
  // 2. first pass: N1-point DFTs over j1 for every (line, j2)
  {
    const int t = threadIdx.x;
    const bool act = t < L * N2;
    const int j2 = t % N2, l = t / N2;
    double2 v[N1];
    if (act) {
      const double *r = R + LS * l;
#pragma unroll
      for (int j1 = 0; j1 < N1; j1++) {
        const int k = N2 * j1 + j2;
        if (!INV) {
          v[j1] = make_double2(r[2 * k], r[2 * k + 1]);
        } else {
          // X[k] and conj X[h-k] from the halfcomplex line
          const double2 xk = (k == 0) ? make_double2(r[0], 0.0) : make_double2(r[k], r[n - k]);
          const double2 xc = (k == 0) ? make_double2(r[h], 0.0) : make_double2(r[h - k], -r[h + k]);
          const double2 w = tw[k];  // conj of w^k
          const double dx = xk.x - xc.x, dy = xk.y - xc.y;
          const double bx = dx * w.x + dy * w.y, by = dy * w.x - dx * w.y;  // (dx + i dy) e^{+2 pi i k/n}
          v[j1] = make_double2(xk.x + xc.x - by, xk.y + xc.y + bx);        // A + i B
        }
      }
      fft_reg<N1>(v, tw, 2 * N2, INV);
#pragma unroll
      for (int k1 = 1; k1 < N1; k1++) {
        double2 w = tw[(2 * j2 * k1) & (n - 1)];
        if (INV) w.y = -w.y;
        double2 y = v[k1];
        v[k1] = make_double2(y.x * w.x - y.y * w.y, y.x * w.y + y.y * w.x);
      }
    }
    __syncthreads();
    if (act) {
      double2 *T = S + (LS / 2) * l;
#pragma unroll
      for (int k1 = 0; k1 < N1; k1++) T[k1 * N2 + (j2 ^ (k1 & (N2 - 1)))] = v[k1];
    }
    __syncthreads();
  }
  // 3. second pass: N2-point DFTs over j2 for every (line, k1)
  {
    const int t = threadIdx.x;
    const bool act = t < L * N1;
    const int k1 = t % N1, l = t / N1;
    double2 v[N2];
    if (act) {
      const double2 *T = S + (LS / 2) * l;
#pragma unroll
      for (int j2 = 0; j2 < N2; j2++) v[j2] = T[k1 * N2 + (j2 ^ (k1 & (N2 - 1)))];
      fft_reg<N2>(v, tw, 2 * N1, INV);
    }
    __syncthreads();
    if (act) {
      double2 *Z = S + (LS / 2) * l;
#pragma unroll
      for (int k2 = 0; k2 < N2; k2++) Z[k1 + N1 * k2] = v[k2];
    }
    __syncthreads();
  }
  // 4. post-mapping and store
#pragma unroll
  for (int it = 0; it < IT; it++) {
    const int e = threadIdx.x + it * NT;
    const int l = a.coal ? e % L : e / n, q = a.coal ? e / L : e % n;
    if (e < L * n && oout[sl][l] >= 0) {
      const double2 *Z = S + (LS / 2) * l;
      double y;
      if (!INV) {
        const int k = q <= h ? q : n - q;
        const double2 zk = Z[k & (h - 1)], zc = Z[(h - k) & (h - 1)];  // Z[k], Z[h-k] (mod h)
        const double ex = 0.5 * (zk.x + zc.x), ey = 0.5 * (zk.y - zc.y);
        const double ox = 0.5 * (zk.y + zc.y), oy = -0.5 * (zk.x - zc.x);
        const double2 w = tw[k];
        const double xr = ex + (w.x * ox - w.y * oy), xi = ey + (w.x * oy + w.y * ox);
        y = q <= h ? xr : xi;
      } else {
        const double2 z = Z[q >> 1];
        y = (q & 1) ? z.y : z.x;
      }
      a.out[q * a.oes + oout[sl][l]] = y * a.scale;
    }
  }
  }
}

// Orders this wavefront's LDS accesses before and after it: a wavefront's
// LDS instructions execute in issue order, so lanes of one wavefront see each
// other's writes without a workgroup barrier once the compiler keeps the
// order (the fences at wavefront scope and the scheduling barrier).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The same transform with each wavefront owning LW = L/4 of the workgroup's
// L lines through both passes, so the passes need no workgroup barrier.
// Contiguous lines (stride 1) are also loaded and stored by their own
// wavefront (one barrier in all, for the line offsets); strided lines are
// loaded and stored by the whole workgroup, 8L contiguous bytes per row,
// with a barrier after the load and one before the store.  Compiled for
// four waves per SIMD (at most 128 VGPRs, a few spilled), which measured
// faster than three without spills.
template <int N1, int N2, bool INV>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(4))) void k_rfft_w(Args a) {
  extern __shared__ double2 S[];
  constexpr int h = N1 * N2, n = 2 * h;
  constexpr int L = rfft_lines(N1, N2), LW = L / 4;
  static_assert(NT == 256 && L % 4 == 0 && LW * N1 <= 64 && LW * N2 <= 64 && (L * n) % NT == 0,
                "one pass task per lane of the owning wavefront");
  constexpr int IT = L * n / NT;  // = LW * n / 64
  const double2 *tw = reinterpret_cast<const double2 *>(a.tab);
  constexpr int LS = n + RPAD;
  double *R = reinterpret_cast<double *>(S);
  __shared__ long oin[L], oout[L];
  const int wv = threadIdx.x >> 6, ln = threadIdx.x & 63;
  for (int t = threadIdx.x; t < L; t += NT) {
    const int line = blockIdx.x * L + t;
    long oi = -1, oo = -1;
    if (line < a.nlines) {
      const int b1 = line % a.nb1, b2 = line / a.nb1;
      oi = b1 * a.is1 + b2 * a.is2;
      oo = b1 * a.os1 + b2 * a.os2;
    }
    oin[t] = oi;
    oout[t] = oo;
  }
  __syncthreads();
  const bool coal = a.coal;
  // element it of this thread: line l, position q
  auto elem = [&](int it, int &l, int &q) {
    if (coal) {
      const int e = threadIdx.x + it * NT;
      l = e % L;
      q = e / L;
    } else {
      const int e = ln + it * 64;
      l = wv * LW + e / n;
      q = e % n;
    }
  };
  // 1. raw lines into LDS
  {
    double xr[IT];
#pragma unroll
    for (int it = 0; it < IT; it++) {
      int l, q;
      elem(it, l, q);
      xr[it] = oin[l] >= 0 ? a.in[q * a.ies + oin[l]] : 0.0;
    }
#pragma unroll
    for (int it = 0; it < IT; it++) {
      int l, q;
      elem(it, l, q);
      R[LS * l + q] = xr[it];
    }
  }
  if (coal) __syncthreads();
  else wave_sync();
  // 2. first pass over this wavefront's lines
  {
    const bool act = ln < LW * N2;
    const int j2 = ln % N2, l = wv * LW + ln / N2;
    double2 v[N1];
    if (act) rfft_pass1<N1, N2, INV>(R + LS * l, j2, tw, v);
    wave_sync();
    if (act) rfft_put1<N1, N2>(S + (LS / 2) * l, j2, v);
    wave_sync();
  }
  // 3. second pass
  {
    const bool act = ln < LW * N1;
    const int k1 = ln % N1, l = wv * LW + ln / N1;
    double2 v[N2];
    if (act) rfft_pass2<N1, N2, INV>(S + (LS / 2) * l, k1, tw, v);
    wave_sync();
    if (act) rfft_put2<N1, N2>(S + (LS / 2) * l, k1, v);
    wave_sync();
  }
  if (coal) __syncthreads();
  // 4. post-mapping and store
#pragma unroll
  for (int it = 0; it < IT; it++) {
    int l, q;
    elem(it, l, q);
    if (oout[l] >= 0) a.out[q * a.oes + oout[l]] = rfft_out<N1, N2, INV>(S + (LS / 2) * l, q, tw) * a.scale;
  }
}

__global__ __launch_bounds__(NT) void k_r2r_dense(Args a) {
  extern __shared__ double2 S[];
  double *R = reinterpret_cast<double *>(S);
  const int n = a.n, L = a.L, line0 = blockIdx.x * L;
  for (int e = threadIdx.x; e < L * n; e += NT) {
    int l, q;
    elem_of(a, e, l, q);
    int line = line0 + l;
    if (line >= a.nlines) continue;
    int b1 = line % a.nb1, b2 = line / a.nb1;
    R[l * n + q] = a.in[q * a.ies + b1 * a.is1 + b2 * a.is2];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < L * n; e += NT) {
    int l, q;
    elem_of(a, e, l, q);
    int line = line0 + l;
    if (line >= a.nlines) continue;
    const double *row = a.tab + (long)q * n;
    const double *x = R + l * n;
    double y = 0.0;
    for (int j = 0; j < n; j++) y += row[j] * x[j];
    int b1 = line % a.nb1, b2 = line / a.nb1;
    a.out[q * a.oes + b1 * a.os1 + b2 * a.os2] = y * a.scale;
  }
}

// n = N1 * N2 with both factors as even as possible
void fft_split(int n, int &n1, int &n2) {
  int lg = 0;
  while ((1 << lg) < n) lg++;
  n2 = 1 << (lg / 2);
  n1 = n / n2;
}

template <int N1, int N2, int LS>
void launch_sized_l(int kind, dim3 grid, size_t lds, hipStream_t s, const Args &a) {
  switch (kind) {
  case K_R2HC: hipLaunchKernelGGL((k_r2r_fft<N1, N2, K_R2HC, LS>), grid, dim3(NT), lds, s, a); break;
  case K_HC2R: hipLaunchKernelGGL((k_r2r_fft<N1, N2, K_HC2R, LS>), grid, dim3(NT), lds, s, a); break;
  case K_REDFT10: hipLaunchKernelGGL((k_r2r_fft<N1, N2, K_REDFT10, LS>), grid, dim3(NT), lds, s, a); break;
  case K_REDFT01: hipLaunchKernelGGL((k_r2r_fft<N1, N2, K_REDFT01, LS>), grid, dim3(NT), lds, s, a); break;
  case K_RODFT10: hipLaunchKernelGGL((k_r2r_fft<N1, N2, K_RODFT10, LS>), grid, dim3(NT), lds, s, a); break;
  case K_RODFT01: hipLaunchKernelGGL((k_r2r_fft<N1, N2, K_RODFT01, LS>), grid, dim3(NT), lds, s, a); break;
  default: throw std::runtime_error("r2r: kind has no FFT path");
  }
}

template <int N1, int N2>
void launch_sized(int kind, int lshift, dim3 grid, size_t lds, hipStream_t s, const Args &a) {
  switch (lshift) {
  case 0: launch_sized_l<N1, N2, 0>(kind, grid, lds, s, a); break;
  case 1: launch_sized_l<N1, N2, 1>(kind, grid, lds, s, a); break;
  case 2: launch_sized_l<N1, N2, 2>(kind, grid, lds, s, a); break;
  default: throw std::runtime_error("r2r: bad tile choice");
  }
}

void launch_fft_kind(int kind, int n, int ls, dim3 grid, size_t lds, hipStream_t s, const Args &a) {
  switch (n) {
  case 4: launch_sized<2, 2>(kind, ls, grid, lds, s, a); break;
  case 8: launch_sized<4, 2>(kind, ls, grid, lds, s, a); break;
  case 16: launch_sized<4, 4>(kind, ls, grid, lds, s, a); break;
  case 32: launch_sized<8, 4>(kind, ls, grid, lds, s, a); break;
  case 64: launch_sized<8, 8>(kind, ls, grid, lds, s, a); break;
  case 128: launch_sized<16, 8>(kind, ls, grid, lds, s, a); break;
  case 256: launch_sized<16, 16>(kind, ls, grid, lds, s, a); break;
  case 512: launch_sized<32, 16>(kind, ls, grid, lds, s, a); break;
  case 1024: launch_sized<32, 32>(kind, ls, grid, lds, s, a); break;
  default: throw std::runtime_error("r2r: no FFT path for length " + std::to_string(n));
  }
}

// tile 0..2: 2^-tile of the lines per workgroup; 3 / 4: full tiles, 2 / 4 per workgroup;
// 5: full tiles, wavefront-local passes (k_rfft_w); 6: full tiles, swapped split;
// 7: wavefront-local passes, swapped split
int rfft_tpb(int tile) { return tile == 3 ? 2 : (tile == 4 ? 4 : 1); }

template <int N1, int N2>
void launch_rfft_sized(bool inv, int tile, dim3 grid, size_t lds, hipStream_t s, const Args &a) {
#define SNAC_RF(I, LS, T) hipLaunchKernelGGL((k_rfft<N1, N2, I, LS, T>), grid, dim3(NT), lds, s, a)
  if (tile == 0) { if (inv) SNAC_RF(true, 0, 1); else SNAC_RF(false, 0, 1); }
  else if (tile == 1) { if (inv) SNAC_RF(true, 1, 1); else SNAC_RF(false, 1, 1); }
  else if (tile == 2) { if (inv) SNAC_RF(true, 2, 1); else SNAC_RF(false, 2, 1); }
  else if (tile == 3) { if (inv) SNAC_RF(true, 0, 2); else SNAC_RF(false, 0, 2); }
  else if (tile == 4) { if (inv) SNAC_RF(true, 0, 4); else SNAC_RF(false, 0, 4); }
  else if (tile == 5) {
    if (inv) hipLaunchKernelGGL((k_rfft_w<N1, N2, true>), grid, dim3(NT), lds, s, a);
    else hipLaunchKernelGGL((k_rfft_w<N1, N2, false>), grid, dim3(NT), lds, s, a);
  }
  else throw std::runtime_error("r2r: bad tile choice");
#undef SNAC_RF
}

// tiles 6 / 7: full tiles (7: wavefront-local passes) with the h-point FFT
// split the other way round (the shorter pass first, over all of a tile's
// threads)
template <int N1, int N2>
void launch_rfft_swapped(bool inv, bool wave, dim3 grid, size_t lds, hipStream_t s, const Args &a) {
  if (wave) {
    if (inv) hipLaunchKernelGGL((k_rfft_w<N1, N2, true>), grid, dim3(NT), lds, s, a);
    else hipLaunchKernelGGL((k_rfft_w<N1, N2, false>), grid, dim3(NT), lds, s, a);
  } else {
    if (inv) hipLaunchKernelGGL((k_rfft<N1, N2, true, 0, 1>), grid, dim3(NT), lds, s, a);
    else hipLaunchKernelGGL((k_rfft<N1, N2, false, 0, 1>), grid, dim3(NT), lds, s, a);
  }
}

// h-point split of an n = 2h point real transform
void launch_rfft(bool inv, int n, int ls, dim3 grid, size_t lds, hipStream_t s, const Args &a) {
  const bool sw = ls == 6 || ls == 7, wv = ls == 7;
  switch (n) {
  case 8: launch_rfft_sized<2, 2>(inv, sw ? (wv ? 5 : 0) : ls, grid, lds, s, a); break;
  case 16:
    if (sw) launch_rfft_swapped<2, 4>(inv, wv, grid, lds, s, a);
    else launch_rfft_sized<4, 2>(inv, ls, grid, lds, s, a);
    break;
  case 32: launch_rfft_sized<4, 4>(inv, sw ? (wv ? 5 : 0) : ls, grid, lds, s, a); break;
  case 64:
    if (sw) launch_rfft_swapped<4, 8>(inv, wv, grid, lds, s, a);
    else launch_rfft_sized<8, 4>(inv, ls, grid, lds, s, a);
    break;
  case 128: launch_rfft_sized<8, 8>(inv, sw ? (wv ? 5 : 0) : ls, grid, lds, s, a); break;
  case 256:
    if (sw) launch_rfft_swapped<8, 16>(inv, wv, grid, lds, s, a);
    else launch_rfft_sized<16, 8>(inv, ls, grid, lds, s, a);
    break;
  case 512: launch_rfft_sized<16, 16>(inv, sw ? (wv ? 5 : 0) : ls, grid, lds, s, a); break;
  case 1024:
    if (sw) launch_rfft_swapped<16, 32>(inv, wv, grid, lds, s, a);
    else launch_rfft_sized<32, 16>(inv, ls, grid, lds, s, a);
    break;
  default: throw std::runtime_error("r2r: no real-FFT path for length " + std::to_string(n));
  }
}

std::mutex g_mu;
std::map<std::pair<int, int>, std::pair<R2RPlan, double *>> g_plans;

bool fft_kind(int kind) {
  return kind == K_R2HC || kind == K_HC2R || kind == K_REDFT10 || kind == K_REDFT01 || kind == K_RODFT10 ||
         kind == K_RODFT01;
}

}  // namespace

double r2r_coef(int kind, int n, int k, int j) {
  const double pi = std::acos(-1.0);
  switch (kind) {
  case K_R2HC:
    if (k <= n / 2) return std::cos(2.0 * pi * (double)j * k / n);
    return -std::sin(2.0 * pi * (double)j * (n - k) / n);
  case K_HC2R:
    if (j == 0) return 1.0;
    if (n % 2 == 0 && j == n / 2) return (k % 2) ? -1.0 : 1.0;
    if (j < (n + 1) / 2) return 2.0 * std::cos(2.0 * pi * (double)j * k / n);
    return -2.0 * std::sin(2.0 * pi * (double)(n - j) * k / n);
  case K_REDFT10: return 2.0 * std::cos(pi * (j + 0.5) * k / n);
  case K_REDFT01: return j == 0 ? 1.0 : 2.0 * std::cos(pi * j * (k + 0.5) / n);
  case K_RODFT10: return 2.0 * std::sin(pi * (j + 0.5) * (k + 1) / n);
  case K_RODFT01: return j == n - 1 ? ((k % 2) ? -1.0 : 1.0) : 2.0 * std::sin(pi * (j + 1) * (k + 0.5) / n);
  case K_REDFT11: return 2.0 * std::cos(pi * (j + 0.5) * (k + 0.5) / n);
  case K_RODFT11: return 2.0 * std::sin(pi * (j + 0.5) * (k + 0.5) / n);
  // DST-I: the transform of a staggered Dirichlet-Dirichlet axis (fft.f90:130-133)
  case K_RODFT00: return 2.0 * std::sin(pi * (j + 1) * (k + 1) / (n + 1));
  default: throw std::runtime_error("unsupported r2r kind " + std::to_string(kind));
  }
}

int r2r_tile_choices(const R2RPlan &pl) {
  if (!pl.fft) return 1;
  return (pl.kind == K_R2HC || pl.kind == K_HC2R) && pl.n >= 8 ? 8 : 3;
}

R2RPlan r2r_plan(int kind, int n) {
  std::lock_guard<std::mutex> lk(g_mu);
  auto key = std::make_pair(kind, n);
  auto it = g_plans.find(key);
  if (it != g_plans.end()) return it->second.first;
  if (n < 1) throw std::runtime_error("r2r: invalid length");
  R2RPlan pl;
  pl.n = n;
  pl.kind = kind;
  pl.fft = fft_kind(kind) && n >= 4 && (n & (n - 1)) == 0 && n <= 1024;
  std::vector<double> h;
  const double pi = std::acos(-1.0);
  if (pl.fft) {
    h.resize(4 * (size_t)n);
    for (int m = 0; m < n; m++) {
      h[2 * m] = std::cos(-2.0 * pi * m / n);
      h[2 * m + 1] = std::sin(-2.0 * pi * m / n);
    }
    for (int k = 0; k < n; k++) {
      h[2 * n + 2 * k] = std::cos(pi * k / (2.0 * n));
      h[2 * n + 2 * k + 1] = std::sin(pi * k / (2.0 * n));
    }
  } else {
    if (n > 2048) throw std::runtime_error("r2r: length " + std::to_string(n) + " needs the FFT path (power of two)");
    h.resize((size_t)n * n);
    for (int k = 0; k < n; k++)
      for (int j = 0; j < n; j++) h[(size_t)k * n + j] = r2r_coef(kind, n, k, j);
  }
  double *d = nullptr;
  SNAC_HIP(hipMalloc(&d, h.size() * sizeof(double)));
  SNAC_HIP(hipMemcpy(d, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
  pl.tab = d;
  g_plans[key] = {pl, d};
  return pl;
}

void launch_r2r(hipStream_t s, const R2RPlan &pl, const double *in, long ies, long is1, long is2, double *out,
                long oes, long os1, long os2, int nb1, int nb2, double scale) {
  Args a;
  a.in = in; a.out = out;
  a.ies = ies; a.is1 = is1; a.is2 = is2; a.oes = oes; a.os1 = os1; a.os2 = os2;
  a.n = pl.n;
  a.logn = 0;
  while ((1 << a.logn) < pl.n) a.logn++;
  a.nb1 = nb1;
  a.nlines = nb1 * nb2;
  a.coal = (ies != 1) ? 1 : 0;
  if (a.nlines <= 0) return;
  if (pl.fft && (pl.kind == K_R2HC || pl.kind == K_HC2R) && pl.n >= 8) {
    int n1 = 0, n2 = 0;
    fft_split(pl.n / 2, n1, n2);
    const int ls = pl.lshift <= 2 ? pl.lshift : 0, tpb = rfft_tpb(pl.lshift);
    const int L = std::max(1, rfft_lines(n1, n2) >> ls);  // the kernel's compile-time line count
    a.L = L;
    a.scale = scale;
    a.tab = pl.tab;
    const int ntile = (a.nlines + L - 1) / L;
    dim3 grid((unsigned)((ntile + tpb - 1) / tpb));
    size_t lds = (size_t)L * (pl.n + RPAD) * sizeof(double);
    launch_rfft(pl.kind == K_HC2R, pl.n, pl.lshift, grid, lds, s, a);
  } else if (pl.fft) {
    int n1 = 0, n2 = 0;
    fft_split(pl.n, n1, n2);
    const int ls = pl.lshift <= 2 ? pl.lshift : 0;
    const int L = std::max(1, fft_lines(n1, n2) >> ls);  // the kernel's compile-time line count
    a.L = L;
    a.scale = scale;
    a.tab = pl.tab;
    dim3 grid((unsigned)((a.nlines + L - 1) / L));
    size_t lds = (size_t)L * pl.n * sizeof(double2);
    launch_fft_kind(pl.kind, pl.n, ls, grid, lds, s, a);
  } else {
    int L = std::max(1, std::min(2 * LDS_COMPLEX / pl.n, a.nlines));
    a.L = L;
    a.scale = scale;
    a.tab = pl.tab;
    dim3 grid((a.nlines + L - 1) / L);
    size_t lds = (size_t)L * pl.n * sizeof(double);
    hipLaunchKernelGGL(k_r2r_dense, grid, dim3(NT), lds, s, a);
  }
  SNAC_HIP(hipGetLastError());
}

}  // namespace snac
