// Batched real-to-real transforms along one axis, the kinds FFTW offers and
// SNaC selects by boundary condition (fft.f90:96-149): R2HC/HC2R (periodic),
// REDFT10/REDFT01 (Neumann-Neumann), RODFT10/RODFT01 (Dirichlet-Dirichlet);
// REDFT11/RODFT11 (mixed) and any length that is not a power of two go
// through a dense matrix path.  Unnormalised, as FFTW's; callers scale.
//
// FFT path: a workgroup owns L whole lines in LDS as complex values, the
// n-point sequence of the kind's pre-mapping (Makhoul's reordering for the
// cosine/sine kinds, halfcomplex unpacking for HC2R) is written straight
// from global memory into bit-reversed slots, radix-2 DIT stages run in LDS
// and the kind's post-mapping is applied while storing.  Loads and stores
// walk the batch index with unit stride when the line stride is not 1, so
// every row is read as one coalesced segment.
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

__device__ __forceinline__ int brev(int x, int logn) { return (int)(__brev((unsigned)x) >> (32 - logn)); }

__device__ __forceinline__ void elem_of(const Args &a, int e, int &line, int &q) {
  if (a.coal) { line = e % a.L; q = e / a.L; }
  else { q = e % a.n; line = e / a.n; }
}

template <int KIND>
__global__ __launch_bounds__(NT) void k_r2r_fft(Args a) {
  extern __shared__ double2 S[];
  const int n = a.n, logn = a.logn, L = a.L;
  const int line0 = blockIdx.x * L;
  const double2 *tw = reinterpret_cast<const double2 *>(a.tab);          // e^{-2 pi i j/n}, j < n/2
  const double2 *tk = reinterpret_cast<const double2 *>(a.tab + n);      // e^{-i pi k/(2n)}, k < n
  // ---- pre-mapping: global -> bit-reversed complex slots ----
  for (int e = threadIdx.x; e < L * n; e += NT) {
    int l, q;
    elem_of(a, e, l, q);
    int line = line0 + l;
    if (line >= a.nlines) continue;
    int b1 = line % a.nb1, b2 = line / a.nb1;
    double x = a.in[q * a.ies + b1 * a.is1 + b2 * a.is2];
    double2 *s = S + l * n;
    if (KIND == K_R2HC) {
      s[brev(q, logn)] = make_double2(x, 0.0);
    } else if (KIND == K_HC2R) {
      const int h = n >> 1;
      if (q == 0 || q == h) s[brev(q, logn)] = make_double2(x, 0.0);
      else if (q < h) { s[brev(q, logn)].x = x; s[brev(n - q, logn)].x = x; }
      else { s[brev(n - q, logn)].y = x; s[brev(q, logn)].y = -x; }
    } else if (KIND == K_REDFT10 || KIND == K_RODFT10) {
      if (KIND == K_RODFT10 && (q & 1)) x = -x;
      int slot = (q & 1) ? n - 1 - (q >> 1) : (q >> 1);
      s[brev(slot, logn)] = make_double2(x, 0.0);
    } else {  // REDFT01 / RODFT01: slot k = e^{i pi k/(2n)} (x_k - i x_{n-k}); first term
      int k = (KIND == K_RODFT01) ? n - 1 - q : q;
      double2 t = tk[k];  // (cos, -sin)
      s[brev(k, logn)] = make_double2(t.x * x, -t.y * x);
    }
  }
  __syncthreads();
  if (KIND == K_REDFT01 || KIND == K_RODFT01) {  // second term into slot n-k
    for (int e = threadIdx.x; e < L * n; e += NT) {
      int l, q;
      elem_of(a, e, l, q);
      int line = line0 + l;
      if (line >= a.nlines) continue;
      int k = (KIND == K_RODFT01) ? n - 1 - q : q;
      if (k == 0) continue;
      int b1 = line % a.nb1, b2 = line / a.nb1;
      double x = a.in[q * a.ies + b1 * a.is1 + b2 * a.is2];
      int m = n - k;
      double2 t = tk[m];  // e^{i pi m/(2n)} * (-i x) = (sin x, -cos x)
      double2 *s = S + l * n;
      double2 v = s[brev(m, logn)];
      v.x += -t.y * x;
      v.y += -t.x * x;
      s[brev(m, logn)] = v;
    }
    __syncthreads();
  }
  // ---- radix-2 DIT stages ----
  const bool inv = (KIND == K_HC2R || KIND == K_REDFT01 || KIND == K_RODFT01);
  const int hn = n >> 1;
  for (int ls = 1; ls <= logn; ls++) {
    const int half = 1 << (ls - 1);
    const int tshift = logn - ls;
    for (int b = threadIdx.x; b < L * hn; b += NT) {
      int l = b / hn, r = b - l * hn;
      int grp = r >> (ls - 1), jj = r & (half - 1);
      int i0 = (grp << ls) + jj, i1 = i0 + half;
      double2 w = tw[jj << tshift];
      if (inv) w.y = -w.y;
      double2 *s = S + l * n;
      double2 x0 = s[i0], x1 = s[i1];
      double2 c = make_double2(x1.x * w.x - x1.y * w.y, x1.x * w.y + x1.y * w.x);
      s[i0] = make_double2(x0.x + c.x, x0.y + c.y);
      s[i1] = make_double2(x0.x - c.x, x0.y - c.y);
    }
    __syncthreads();
  }
  // ---- post-mapping and store ----
  for (int e = threadIdx.x; e < L * n; e += NT) {
    int l, q;
    elem_of(a, e, l, q);
    int line = line0 + l;
    if (line >= a.nlines) continue;
    const double2 *s = S + l * n;
    double y;
    if (KIND == K_R2HC) {
      y = (q <= hn) ? s[q].x : s[n - q].y;
    } else if (KIND == K_HC2R) {
      y = s[q].x;
    } else if (KIND == K_REDFT10 || KIND == K_RODFT10) {
      int k = (KIND == K_RODFT10) ? n - 1 - q : q;
      double2 t = tk[k], z = s[k];
      y = 2.0 * (z.x * t.x - z.y * t.y);
    } else {
      int slot = (q & 1) ? n - 1 - (q >> 1) : (q >> 1);
      y = s[slot].x;
      if (KIND == K_RODFT01 && (q & 1)) y = -y;
    }
    int b1 = line % a.nb1, b2 = line / a.nb1;
    a.out[q * a.oes + b1 * a.os1 + b2 * a.os2] = y * a.scale;
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
  default: throw std::runtime_error("unsupported r2r kind " + std::to_string(kind));
  }
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
  pl.fft = fft_kind(kind) && n >= 4 && (n & (n - 1)) == 0 && n <= LDS_COMPLEX;
  std::vector<double> h;
  const double pi = std::acos(-1.0);
  if (pl.fft) {
    h.resize(3 * (size_t)n);
    for (int j = 0; j < n / 2; j++) {
      h[2 * j] = std::cos(-2.0 * pi * j / n);
      h[2 * j + 1] = std::sin(-2.0 * pi * j / n);
    }
    for (int k = 0; k < n; k++) {
      h[n + 2 * k] = std::cos(-pi * k / (2.0 * n));
      h[n + 2 * k + 1] = std::sin(-pi * k / (2.0 * n));
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
  int per = pl.fft ? LDS_COMPLEX : 2 * LDS_COMPLEX;  // lines per workgroup from the LDS budget
  int L = per / pl.n;
  if (L < 1) L = 1;
  if (L > a.nlines) L = a.nlines;
  a.L = L;
  a.scale = scale;
  a.tab = pl.tab;
  if (a.nlines <= 0) return;
  dim3 grid((a.nlines + L - 1) / L);
  if (pl.fft) {
    size_t lds = (size_t)L * pl.n * sizeof(double2);
    switch (pl.kind) {
    case K_R2HC: hipLaunchKernelGGL(k_r2r_fft<K_R2HC>, grid, dim3(NT), lds, s, a); break;
    case K_HC2R: hipLaunchKernelGGL(k_r2r_fft<K_HC2R>, grid, dim3(NT), lds, s, a); break;
    case K_REDFT10: hipLaunchKernelGGL(k_r2r_fft<K_REDFT10>, grid, dim3(NT), lds, s, a); break;
    case K_REDFT01: hipLaunchKernelGGL(k_r2r_fft<K_REDFT01>, grid, dim3(NT), lds, s, a); break;
    case K_RODFT10: hipLaunchKernelGGL(k_r2r_fft<K_RODFT10>, grid, dim3(NT), lds, s, a); break;
    case K_RODFT01: hipLaunchKernelGGL(k_r2r_fft<K_RODFT01>, grid, dim3(NT), lds, s, a); break;
    default: throw std::runtime_error("r2r: bad fft kind");
    }
  } else {
    size_t lds = (size_t)L * pl.n * sizeof(double);
    hipLaunchKernelGGL(k_r2r_dense, grid, dim3(NT), lds, s, a);
  }
  SNAC_HIP(hipGetLastError());
}

}  // namespace snac
