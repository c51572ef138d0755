// The learned preconditioner's network (net.hpp): BF16 MFMA convolutions
// with FP32 accumulation (forward, data gradient, weight gradient), the
// pooling / upsampling / ReLU passes, the loss and Adam.
//
// Convolutions are implicit GEMMs on v_mfma_f32_16x16x16_bf16: for the
// forward pass one wavefront owns 16 pixels of one image row and every
// output channel: per tap and 16-channel slice of the input, A = 16 pixels x
// 16 input channels (each lane converts one float4 of its pixel's channels
// to BF16), B = 16 input channels x 16 output channels of the BF16 weights,
// D accumulates in FP32.  The data gradient is the same kernel on the
// flipped, channel-transposed weights; the weight gradient reduces
// pixels x pixels GEMMs (A = gradient, B = shifted input) per workgroup and
// adds them into the FP32 gradient with vector atomics.
#include "common.hpp"
#include "net.hpp"

#include <cmath>
#include <cstring>

namespace snac {
namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__host__ __device__ inline short bf16_bits(float f) {
  unsigned u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) return (short)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}

// out[b][y][x][co] = act(bias[co] + sum_{tap,ci} in[b][y+dy][x+dx][ci] * w[tap][co][ci])
// zero padding; a 256-thread workgroup covers 64 pixels of row y of image b
template <int CIN, int COUT, bool RELU>
__global__ __launch_bounds__(256) void k_conv(int H, int W, const float *__restrict__ in,
                                              const short *__restrict__ w, const float *__restrict__ bias,
                                              float *__restrict__ out) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int x0 = (blockIdx.x * 4 + wv) * 16;
  const int y = blockIdx.y, b = blockIdx.z;
  const int px = x0 + (lane & 15), kq = (lane >> 4) * 4;
  f32x4 acc[COUT / 16];
#pragma unroll
  for (int t = 0; t < COUT / 16; t++) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
  const float *img = in + (size_t)b * H * W * CIN;
  for (int tap = 0; tap < 9; tap++) {
    const int yy = y + tap / 3 - 1, xx = px + tap % 3 - 1;
    const bool ok = px < W && yy >= 0 && yy < H && xx >= 0 && xx < W;
    const float *src = img + ((size_t)(ok ? yy : 0) * W + (ok ? xx : 0)) * CIN + kq;
#pragma unroll
    for (int c0 = 0; c0 < CIN; c0 += 16) {
      bf16x4 a = bf16x4{0, 0, 0, 0};
      if (ok) {
        const float4 v = *(const float4 *)(src + c0);
        a = bf16x4{bf16_bits(v.x), bf16_bits(v.y), bf16_bits(v.z), bf16_bits(v.w)};
      }
#pragma unroll
      for (int t = 0; t < COUT / 16; t++) {
        const bf16x4 bv = *(const bf16x4 *)(w + ((size_t)tap * COUT + t * 16 + (lane & 15)) * CIN + c0 + kq);
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bv, acc[t], 0, 0, 0);
      }
    }
  }
  // D: lane holds pixels x0 + kq + j of output channel t*16 + (lane & 15)
#pragma unroll
  for (int t = 0; t < COUT / 16; t++) {
    const int co = t * 16 + (lane & 15);
    const float bb = bias ? bias[co] : 0.f;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int x = x0 + kq + j;
      if (x < W) {
        float v = acc[t][j] + bb;
        if (RELU) v = fmaxf(v, 0.f);
        out[(((size_t)b * H + y) * W + x) * COUT + co] = v;
      }
    }
  }
}

// The same convolution with the input staged in LDS: a 256-thread
// workgroup computes 4 output rows x 64 pixels (one row per wavefront, four
// 16-pixel MFMA tiles each); the 6 x 66 input pixels it reads (one-pixel
// halo, zero outside the image) are converted to BF16 once into LDS, each
// pixel padded by 8 BF16 so consecutive pixels start on different banks.
// Every input pixel is then fetched from memory about 1.5 times instead of
// once per tap.
template <int CIN, int COUT, bool RELU>
__global__ __launch_bounds__(256) void k_conv_lds(int H, int W, const float *__restrict__ in,
                                                  const short *__restrict__ w, const float *__restrict__ bias,
                                                  float *__restrict__ out) {
  constexpr int TC = 66, TR = 6, PS = CIN + 8;  // tile columns, rows, pixel stride (BF16)
  __shared__ short tile[TR * TC * PS];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int x0 = blockIdx.x * 64, y0 = blockIdx.y * 4, b = blockIdx.z;
  const float *img = in + (size_t)b * H * W * CIN;
  for (int q = threadIdx.x; q < TR * TC * (CIN / 4); q += 256) {
    const int cq = q % (CIN / 4), pc = q / (CIN / 4);
    const int c = pc % TC, r = pc / TC;
    const int yy = y0 - 1 + r, xx = x0 - 1 + c;
    bf16x4 v = bf16x4{0, 0, 0, 0};
    if (yy >= 0 && yy < H && xx >= 0 && xx < W) {
      const float4 f = *(const float4 *)(img + ((size_t)yy * W + xx) * CIN + cq * 4);
      v = bf16x4{bf16_bits(f.x), bf16_bits(f.y), bf16_bits(f.z), bf16_bits(f.w)};
    }
    *(bf16x4 *)(tile + pc * PS + cq * 4) = v;
  }
  __syncthreads();
  const int kq = (lane >> 4) * 4, y = y0 + wv;
  f32x4 acc[4][COUT / 16];
#pragma unroll
  for (int s4 = 0; s4 < 4; s4++)
#pragma unroll
    for (int t = 0; t < COUT / 16; t++) acc[s4][t] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int tap = 0; tap < 9; tap++) {
    const int r = wv + tap / 3, cs = (lane & 15) + tap % 3;  // tile row / column of this lane's pixel (sub-tile 0)
#pragma unroll
    for (int c0 = 0; c0 < CIN; c0 += 16) {
      bf16x4 bv[COUT / 16];
#pragma unroll
      for (int t = 0; t < COUT / 16; t++)
        bv[t] = *(const bf16x4 *)(w + ((size_t)tap * COUT + t * 16 + (lane & 15)) * CIN + c0 + kq);
#pragma unroll
      for (int s4 = 0; s4 < 4; s4++) {
        const bf16x4 a = *(const bf16x4 *)(tile + (r * TC + cs + s4 * 16) * PS + c0 + kq);
#pragma unroll
        for (int t = 0; t < COUT / 16; t++)
          acc[s4][t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bv[t], acc[s4][t], 0, 0, 0);
      }
    }
  }
  if (y >= H) return;
#pragma unroll
  for (int t = 0; t < COUT / 16; t++) {
    const int co = t * 16 + (lane & 15);
    const float bb = bias ? bias[co] : 0.f;
#pragma unroll
    for (int s4 = 0; s4 < 4; s4++)
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int x = x0 + s4 * 16 + kq + j;
        if (x < W) {
          float v = acc[s4][t][j] + bb;
          if (RELU) v = fmaxf(v, 0.f);
          out[(((size_t)b * H + y) * W + x) * COUT + co] = v;
        }
      }
  }
}

// dw[tap][co][ci] += sum_p g[p][co] * in[p + off(tap)][ci] over pixels p of
// all images; workgroup (pixel range, tap), wavefront = 16 output channels
template <int CIN, int COUT>
__global__ __launch_bounds__(256) void k_conv_wgrad(int B, int H, int W, const float *__restrict__ in,
                                                    const float *__restrict__ g, float *__restrict__ dw, int prange) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int tap = blockIdx.y, dy = tap / 3 - 1, dx = tap % 3 - 1;
  const int kq = (lane >> 4) * 4;
  const long npix = (long)B * H * W;
  const long p0 = (long)blockIdx.x * prange, p1 = p0 + prange < npix ? p0 + prange : npix;
  f32x4 acc[CIN / 16];
#pragma unroll
  for (int t = 0; t < CIN / 16; t++) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int co = wv * 16 + (lane & 15);
  for (long pb = p0; pb < p1; pb += 16) {
    bf16x4 a;
    long src[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const long p = pb + kq + j;
      a[j] = p < p1 ? bf16_bits(g[p * COUT + co]) : (short)0;
      src[j] = -1;
      if (p < p1) {
        const int x = (int)(p % W), y = (int)((p / W) % H);
        const long bi = p / ((long)H * W);
        if (y + dy >= 0 && y + dy < H && x + dx >= 0 && x + dx < W) src[j] = ((bi * H + y + dy) * W + x + dx) * CIN;
      }
    }
#pragma unroll
    for (int c0 = 0; c0 < CIN; c0 += 16) {
      bf16x4 bv;
#pragma unroll
      for (int j = 0; j < 4; j++) bv[j] = src[j] >= 0 ? bf16_bits(in[src[j] + c0 + (lane & 15)]) : (short)0;
      acc[c0 / 16] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bv, acc[c0 / 16], 0, 0, 0);
    }
  }
  // D[i = co][n = ci]: lane holds co = wv*16 + kq + j, ci = t*16 + (lane & 15)
#pragma unroll
  for (int t = 0; t < CIN / 16; t++)
#pragma unroll
    for (int j = 0; j < 4; j++)
      atomicAdd(&dw[((size_t)tap * COUT + wv * 16 + kq + j) * CIN + t * 16 + (lane & 15)], acc[t][j]);
}

// db[co] += sum_p g[p][co]
template <int COUT>
__global__ __launch_bounds__(256) void k_bias_grad(long npix, const float *__restrict__ g, float *__restrict__ db) {
  __shared__ float part[256];
  const int co = threadIdx.x % COUT, r = threadIdx.x / COUT, rows = 256 / COUT;
  float s = 0.f;
  for (long p = (long)blockIdx.x * rows + r; p < npix; p += (long)gridDim.x * rows) s += g[p * COUT + co];
  part[threadIdx.x] = s;
  __syncthreads();
  if (r == 0) {
    for (int q = 1; q < rows; q++) s += part[q * COUT + co];
    atomicAdd(&db[co], s);
  }
}

// 2x2 average pooling (n = pixels of the output, C channels)
__global__ void k_pool(int B, int Ho, int Wo, int C, const float *__restrict__ in, float *__restrict__ out) {
  const long n = (long)B * Ho * Wo * C;
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
    const int c = (int)(q % C);
    const long p = q / C;
    const int x = (int)(p % Wo), y = (int)((p / Wo) % Ho);
    const long b = p / ((long)Ho * Wo);
    const long Wi = 2L * Wo;
    const long base = ((b * 2 * Ho + 2 * y) * Wi + 2 * x) * C + c;
    out[q] = 0.25f * (in[base] + in[base + C] + in[base + Wi * C] + in[base + Wi * C + C]);
  }
}

// g_in (fine) += g_out (coarse) / 4 at each of the four children
__global__ void k_pool_bwd_add(int B, int H, int W, int C, const float *__restrict__ gc, float *__restrict__ gf) {
  const long n = (long)B * H * W * C;
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
    const int c = (int)(q % C);
    const long p = q / C;
    const int x = (int)(p % W), y = (int)((p / W) % H);
    const long b = p / ((long)H * W);
    gf[q] += 0.25f * gc[((b * (H / 2) + y / 2) * (W / 2) + x / 2) * C + c];
  }
}

// out (fine) = coarse(parent) + skip
__global__ void k_up_add(int B, int H, int W, int C, const float *__restrict__ coarse, const float *__restrict__ skip,
                         float *__restrict__ out) {
  const long n = (long)B * H * W * C;
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
    const int c = (int)(q % C);
    const long p = q / C;
    const int x = (int)(p % W), y = (int)((p / W) % H);
    const long b = p / ((long)H * W);
    out[q] = coarse[((b * (H / 2) + y / 2) * (W / 2) + x / 2) * C + c] + skip[q];
  }
}

// g_coarse = sum of the four children's gradients (n = coarse elements)
__global__ void k_up_bwd(int B, int Ho, int Wo, int C, const float *__restrict__ gf, float *__restrict__ gc) {
  const long n = (long)B * Ho * Wo * C;
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
    const int c = (int)(q % C);
    const long p = q / C;
    const int x = (int)(p % Wo), y = (int)((p / Wo) % Ho);
    const long b = p / ((long)Ho * Wo);
    const long Wi = 2L * Wo;
    const long base = ((b * 2 * Ho + 2 * y) * Wi + 2 * x) * C + c;
    gc[q] = gf[base] + gf[base + C] + gf[base + Wi * C] + gf[base + Wi * C + C];
  }
}

__global__ void k_relu_bwd(long n, const float *__restrict__ act, float *__restrict__ g) {
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x)
    if (!(act[q] > 0.f)) g[q] = 0.f;
}

__global__ void k_copy(long n, const float *__restrict__ a, float *__restrict__ b) {
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) b[q] = a[q];
}

// per image: s1 = sum (y - t)^2, s2 = sum t^2 (one workgroup per image)
__global__ __launch_bounds__(256) void k_loss_sums(long plane, const float *__restrict__ y, const float *__restrict__ t,
                                                   float *__restrict__ part) {
  __shared__ float a[256], c[256];
  const int b = blockIdx.x;
  float s1 = 0.f, s2 = 0.f;
  for (long p = threadIdx.x; p < plane; p += 256) {
    const float tv = t[b * plane + p], d = y[(b * plane + p) * 16] - tv;
    s1 += d * d;
    s2 += tv * tv;
  }
  a[threadIdx.x] = s1;
  c[threadIdx.x] = s2;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) { a[threadIdx.x] += a[threadIdx.x + o]; c[threadIdx.x] += c[threadIdx.x + o]; }
    __syncthreads();
  }
  if (threadIdx.x == 0) { part[2 * b] = a[0]; part[2 * b + 1] = c[0]; }
}

// gy (channel 0) = 2 (y - t) / (B |t_b|^2), other channels 0; loss = mean_b s1/s2
__global__ void k_loss_grad(int B, long plane, const float *__restrict__ y, const float *__restrict__ t,
                            const float *__restrict__ part, float *__restrict__ gy, float *__restrict__ loss) {
  const long n = (long)B * plane * 16;
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
    const long p = q / 16;
    const int b = (int)(p / plane);
    const float s2 = part[2 * b + 1] > 0.f ? part[2 * b + 1] : 1.f;
    gy[q] = (q % 16 == 0) ? 2.f * (y[q] - t[p]) / ((float)B * s2) : 0.f;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    float l = 0.f;
    for (int b = 0; b < B; b++) l += part[2 * b] / (part[2 * b + 1] > 0.f ? part[2 * b + 1] : 1.f);
    loss[0] = l / B;
  }
}

__global__ void k_adam(long n, float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, float lr, float c1, float c2) {
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
    const float gq = g[q];
    const float mq = b1 * m[q] + (1.f - b1) * gq, vq = b2 * v[q] + (1.f - b2) * gq * gq;
    m[q] = mq;
    v[q] = vq;
    p[q] -= lr * (mq * c1) / (sqrtf(vq * c2) + eps);
  }
}

// BF16 weights [tap][co][ci] and the flipped, transposed copy
// wf[tap][ci][co] = w[8 - tap][co][ci] (the data-gradient kernel's weights)
__global__ void k_wviews(int cin, int cout, const float *__restrict__ w32, short *__restrict__ wb,
                         short *__restrict__ wf) {
  const int n = 9 * cin * cout;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
    const int ci = q % cin, co = (q / cin) % cout, tap = q / (cin * cout);
    const short v = bf16_bits(w32[q]);
    wb[q] = v;
    wf[((size_t)(8 - tap) * cin + ci) * cout + co] = v;
  }
}

__global__ void k_expand2(long npix, const float *__restrict__ x2, float *__restrict__ x16) {
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < npix * 16; q += (long)gridDim.x * blockDim.x) {
    const int c = (int)(q % 16);
    x16[q] = c < 2 ? x2[(q / 16) * 2 + c] : 0.f;
  }
}

__global__ void k_channel0(long npix, const float *__restrict__ y16, float *__restrict__ y) {
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < npix; q += (long)gridDim.x * blockDim.x)
    y[q] = y16[q * 16];
}

__global__ void k_net_in(int ng, int na, int ns, int nH, int nW, const float *__restrict__ r,
                         const double *__restrict__ ss, double nact, const double *__restrict__ lam, double ilam,
                         float *__restrict__ x16) {
  const long n = (long)ng * nH * nW;
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
    const int x = (int)(q % nW), y = (int)((q / nW) % nH), g = (int)(q / ((long)nH * nW));
    const double rms = sqrt(ss[2 * g] / nact);
    const float v = (y < na && x < ns && rms > 0.0) ? (float)(r[((long)g * na + y) * ns + x] / rms) : 0.f;
    float4 *o = (float4 *)(x16 + q * 16);
    o[0] = make_float4(v, (float)(lam[g] * ilam), 0.f, 0.f);
    o[1] = o[2] = o[3] = make_float4(0.f, 0.f, 0.f, 0.f);
  }
}

__global__ void k_net_out(int ng, int na, int ns, int nH, int nW, const float *__restrict__ y16,
                          const double *__restrict__ ss, double nact, const unsigned char *__restrict__ act,
                          float *__restrict__ z) {
  const long n = (long)ng * na * ns;
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
    const int x = (int)(q % ns), y = (int)((q / ns) % na), g = (int)(q / ((long)na * ns));
    const double rms = sqrt(ss[2 * g] / nact);
    z[q] = act[(long)y * ns + x] ? (float)(rms * y16[(((long)g * nH + y) * nW + x) * 16]) : 0.f;
  }
}

__global__ void k_net_pair_in(int ng, int na, int ns, int nH, int nW, const double *__restrict__ b,
                              const double *__restrict__ bb, double nact, const double *__restrict__ lam, double ilam,
                              float *__restrict__ x2) {
  const long n = (long)ng * nH * nW;
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
    const int x = (int)(q % nW), y = (int)((q / nW) % nH), g = (int)(q / ((long)nH * nW));
    const double rms = sqrt(bb[2 * g] / nact);
    x2[2 * q] = (y < na && x < ns && rms > 0.0) ? (float)(b[((long)g * na + y) * ns + x] / rms) : 0.f;
    x2[2 * q + 1] = (float)(lam[g] * ilam);
  }
}

__global__ void k_net_pair_out(int ng, int na, int ns, int nH, int nW, const double *__restrict__ xs,
                               const double *__restrict__ bb, double nact, float *__restrict__ t) {
  const long n = (long)ng * nH * nW;
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
    const int x = (int)(q % nW), y = (int)((q / nW) % nH), g = (int)(q / ((long)nH * nW));
    const double rms = sqrt(bb[2 * g] / nact);
    t[q] = (y < na && x < ns && rms > 0.0) ? (float)(xs[((long)g * na + y) * ns + x] / rms) : 0.f;
  }
}

__global__ void k_net_stall(int ng, const double *__restrict__ rr, double *__restrict__ prev,
                            const int *__restrict__ done, double ratio2, int *__restrict__ stall) {
  for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < ng; g += gridDim.x * blockDim.x) {
    if (!done[g] && !(rr[2 * g] <= ratio2 * prev[2 * g])) stall[0] = 1;  // a NaN residual stalls too
    prev[2 * g] = rr[2 * g];
  }
}

unsigned grid_for(long n) { return (unsigned)std::min<long>((n + 255) / 256, 8192); }

// 0: one wavefront per 16 pixels reading global memory per tap; 1: the
// LDS-staged form (default)
int g_conv_variant = 1;

template <int CIN, int COUT>
void conv_fwd(hipStream_t s, int B, int H, int W, const float *in, const short *w, const float *bias, float *out,
              bool relu) {
  if (g_conv_variant == 1) {
    dim3 grid((W + 63) / 64, (H + 3) / 4, B);
    if (relu) hipLaunchKernelGGL((k_conv_lds<CIN, COUT, true>), grid, dim3(256), 0, s, H, W, in, w, bias, out);
    else hipLaunchKernelGGL((k_conv_lds<CIN, COUT, false>), grid, dim3(256), 0, s, H, W, in, w, bias, out);
    return;
  }
  dim3 grid((W + 63) / 64, H, B);
  if (relu) hipLaunchKernelGGL((k_conv<CIN, COUT, true>), grid, dim3(256), 0, s, H, W, in, w, bias, out);
  else hipLaunchKernelGGL((k_conv<CIN, COUT, false>), grid, dim3(256), 0, s, H, W, in, w, bias, out);
}

// dispatch on the layer's (cin, cout); dgrad swaps them
void conv_any(hipStream_t s, int cin, int cout, int B, int H, int W, const float *in, const short *w,
              const float *bias, float *out, bool relu) {
#define SNAC_CONV(CI, CO) \
  if (cin == CI && cout == CO) return conv_fwd<CI, CO>(s, B, H, W, in, w, bias, out, relu);
  SNAC_CONV(16, 16) SNAC_CONV(16, 32) SNAC_CONV(32, 16) SNAC_CONV(32, 64) SNAC_CONV(64, 32) SNAC_CONV(64, 64)
#undef SNAC_CONV
  throw std::runtime_error("net: no convolution kernel for these channel counts");
}

void wgrad_any(hipStream_t s, int cin, int cout, int B, int H, int W, const float *in, const float *g, float *dw) {
  const long npix = (long)B * H * W;
  const int prange = 4096;
  dim3 grid((unsigned)((npix + prange - 1) / prange), 9);
#define SNAC_WG(CI, CO)                                                                                            \
  if (cin == CI && cout == CO) {                                                                                   \
    hipLaunchKernelGGL((k_conv_wgrad<CI, CO>), grid, dim3(64 * (CO / 16)), 0, s, B, H, W, in, g, dw, prange);       \
    return;                                                                                                        \
  }
  SNAC_WG(16, 16) SNAC_WG(16, 32) SNAC_WG(32, 64) SNAC_WG(64, 64) SNAC_WG(64, 32) SNAC_WG(32, 16)
#undef SNAC_WG
  throw std::runtime_error("net: no weight-gradient kernel for these channel counts");
}

void bias_grad_any(hipStream_t s, int cout, long npix, const float *g, float *db) {
  const unsigned grid = (unsigned)std::min<long>((npix * cout + 255) / 256, 1024);
  if (cout == 16) hipLaunchKernelGGL(k_bias_grad<16>, dim3(grid), dim3(256), 0, s, npix, g, db);
  else if (cout == 32) hipLaunchKernelGGL(k_bias_grad<32>, dim3(grid), dim3(256), 0, s, npix, g, db);
  else hipLaunchKernelGGL(k_bias_grad<64>, dim3(grid), dim3(256), 0, s, npix, g, db);
}

}  // namespace

void launch_net_in(hipStream_t s, int ng, int na, int ns, int nH, int nW, const float *r, const double *ss,
                   double nact, const double *lam, double lam_ref, float *x16) {
  hipLaunchKernelGGL(k_net_in, dim3(grid_for((long)ng * nH * nW)), dim3(256), 0, s, ng, na, ns, nH, nW, r, ss, nact,
                     lam, 1.0 / lam_ref, x16);
  SNAC_HIP(hipGetLastError());
}

void launch_net_out(hipStream_t s, int ng, int na, int ns, int nH, int nW, const float *y16, const double *ss,
                    double nact, const unsigned char *act, float *z) {
  hipLaunchKernelGGL(k_net_out, dim3(grid_for((long)ng * na * ns)), dim3(256), 0, s, ng, na, ns, nH, nW, y16, ss,
                     nact, act, z);
  SNAC_HIP(hipGetLastError());
}

void launch_net_pair_in(hipStream_t s, int ng, int na, int ns, int nH, int nW, const double *b, const double *bb,
                        double nact, const double *lam, double lam_ref, float *x2) {
  hipLaunchKernelGGL(k_net_pair_in, dim3(grid_for((long)ng * nH * nW)), dim3(256), 0, s, ng, na, ns, nH, nW, b, bb,
                     nact, lam, 1.0 / lam_ref, x2);
  SNAC_HIP(hipGetLastError());
}

void launch_net_pair_out(hipStream_t s, int ng, int na, int ns, int nH, int nW, const double *x, const double *bb,
                         double nact, float *t) {
  hipLaunchKernelGGL(k_net_pair_out, dim3(grid_for((long)ng * nH * nW)), dim3(256), 0, s, ng, na, ns, nH, nW, x, bb,
                     nact, t);
  SNAC_HIP(hipGetLastError());
}

void launch_net_stall(hipStream_t s, int ng, const double *rr, double *prev, const int *done, double ratio2,
                      int *stall) {
  hipLaunchKernelGGL(k_net_stall, dim3((ng + 255) / 256), dim3(256), 0, s, ng, rr, prev, done, ratio2, stall);
  SNAC_HIP(hipGetLastError());
}

void net_set_conv_variant(int v) {
  if (v < 0 || v > 1) throw std::runtime_error("convolution variant must be 0 or 1");
  g_conv_variant = v;
}

float bf16_round(float x) {
  const unsigned short b = (unsigned short)bf16_bits(x);
  const unsigned u = (unsigned)b << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

float *Net::alloc(size_t n) {
  void *p = nullptr;
  SNAC_HIP(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(float)));
  SNAC_HIP(hipMemset(p, 0, std::max<size_t>(n, 1) * sizeof(float)));
  owned_.push_back(p);
  return (float *)p;
}

Net::Net(int H, int W, int maxbatch, unsigned seed) : H_(H), W_(W), mb_(maxbatch) {
  if (H < 8 || W < 8 || H % 8 || W % 8) throw std::runtime_error("net: plane sizes must be multiples of 8");
  if (maxbatch < 1) throw std::runtime_error("net: batch must be positive");
  size_t off = 0;
  for (int l = 0; l < NET_LAYERS; l++) {
    woff_[l] = off;
    off += 9 * (size_t)NET_CIN[l] * NET_COUT[l];
    boff_[l] = off;
    off += NET_COUT[l];
  }
  nparam_ = off;
  p32_ = alloc(nparam_);
  g32_ = alloc(nparam_);
  m_ = alloc(nparam_);
  v_ = alloc(nparam_);
  wb_ = (short *)alloc(nparam_ / 2 + 1);
  wf_ = (short *)alloc(nparam_ / 2 + 1);
  // He-uniform weights over the live channels (layer 0 reads 2 input
  // channels, layer 6 writes 1 output channel), zero biases
  std::vector<float> h(nparam_, 0.f);
  unsigned long long st = 0x9E3779B97F4A7C15ull ^ seed;
  auto rnd = [&]() {
    st ^= st << 13; st ^= st >> 7; st ^= st << 17;
    return (float)((st >> 11) * (1.0 / 9007199254740992.0)) * 2.f - 1.f;
  };
  for (int l = 0; l < NET_LAYERS; l++) {
    const int cin_live = l == 0 ? 2 : NET_CIN[l], cout_live = l == NET_LAYERS - 1 ? 1 : NET_COUT[l];
    const float a = std::sqrt(6.f / (9.f * cin_live)) * (l == NET_LAYERS - 1 ? 0.1f : 1.f);
    for (int tap = 0; tap < 9; tap++)
      for (int co = 0; co < NET_COUT[l]; co++)
        for (int ci = 0; ci < NET_CIN[l]; ci++) {
          const float v = rnd() * a;
          h[woff_[l] + ((size_t)tap * NET_COUT[l] + co) * NET_CIN[l] + ci] = (ci < cin_live && co < cout_live) ? v : 0.f;
        }
  }
  set_params(h.data());
  // activations (NHWC FP32) at each level for the largest batch
  const size_t B = mb_;
  act_[0] = alloc(B * plane(0) * 16);
  for (int l = 0; l < NET_LAYERS - 1; l++) act_[l + 1] = alloc(B * plane(NET_LEVEL[l]) * NET_COUT[l]);
  for (int q = 0; q < 3; q++) {
    // pool_[q]: level q+1 with the channels of e_(q+1); sum_[q]: level q with
    // the channels of the skip e_(q+1)
    const int ch = NET_COUT[q];
    pool_[q] = alloc(B * plane(q + 1) * ch);
    gpool_[q] = alloc(B * plane(q + 1) * ch);
    sum_[q] = alloc(B * plane(q) * ch);
    gsum_[q] = alloc(B * plane(q) * ch);
  }
  for (int l = 0; l < NET_LAYERS - 1; l++) gact_[l + 1] = alloc(B * plane(NET_LEVEL[l]) * NET_COUT[l]);
  y_ = alloc(B * plane(0) * 16);
  gy_ = alloc(B * plane(0) * 16);
  loss_part_ = alloc(2 * B + 2);
}

Net::~Net() {
  for (void *p : owned_) (void)hipFree(p);
}

void Net::refresh_bf16(hipStream_t s) {
  for (int l = 0; l < NET_LAYERS; l++)
    hipLaunchKernelGGL(k_wviews, dim3(grid_for(9L * NET_CIN[l] * NET_COUT[l])), dim3(256), 0, s, NET_CIN[l],
                       NET_COUT[l], p32_ + woff_[l], wb_ + woff_[l], wf_ + woff_[l]);
  SNAC_HIP(hipGetLastError());
}

void Net::get_params(float *host) const {
  SNAC_HIP(hipMemcpy(host, p32_, nparam_ * sizeof(float), hipMemcpyDeviceToHost));
}

void Net::set_params(const float *host) {
  SNAC_HIP(hipMemcpy(p32_, host, nparam_ * sizeof(float), hipMemcpyHostToDevice));
  refresh_bf16(nullptr);
  SNAC_HIP(hipDeviceSynchronize());
}

void Net::grads_to_host(float *host) const {
  SNAC_HIP(hipMemcpy(host, g32_, nparam_ * sizeof(float), hipMemcpyDeviceToHost));
}

void Net::set_input2(hipStream_t s, int B, const float *x2) {
  const long np = (long)B * plane(0);
  hipLaunchKernelGGL(k_expand2, dim3(grid_for(np * 16)), dim3(256), 0, s, np, x2, act_[0]);
  SNAC_HIP(hipGetLastError());
}

void Net::get_output(hipStream_t s, int B, float *y) const {
  const long np = (long)B * plane(0);
  hipLaunchKernelGGL(k_channel0, dim3(grid_for(np)), dim3(256), 0, s, np, y_, y);
  SNAC_HIP(hipGetLastError());
}

void Net::forward(hipStream_t s, int B) {
  if (B < 1 || B > mb_) throw std::runtime_error("net: batch out of range");
  const float *in[NET_LAYERS] = {act_[0], pool_[0], pool_[1], pool_[2], sum_[2], sum_[1], sum_[0]};
  for (int l = 0; l < NET_LAYERS; l++) {
    const int lv = NET_LEVEL[l], h = H_ >> lv, w = W_ >> lv;
    // the inputs of layers 1-3 are pooled, of layers 4-6 upsampled + skip
    if (l >= 1 && l <= 3) {
      const int ch = NET_CIN[l];
      const long n = (long)B * h * w * ch;
      hipLaunchKernelGGL(k_pool, dim3(grid_for(n)), dim3(256), 0, s, B, h, w, ch, act_[l], pool_[l - 1]);
    } else if (l >= 4) {
      const int q = 6 - l;  // sum_[2] for layer 4, ..., sum_[0] for layer 6
      const int ch = NET_CIN[l];
      const long n = (long)B * h * w * ch;
      hipLaunchKernelGGL(k_up_add, dim3(grid_for(n)), dim3(256), 0, s, B, h, w, ch, act_[l], act_[q + 1], sum_[q]);
    }
    float *out = l == NET_LAYERS - 1 ? y_ : act_[l + 1];
    conv_any(s, NET_CIN[l], NET_COUT[l], B, h, w, in[l], wb_ + woff_[l], p32_ + boff_[l], out, l < NET_LAYERS - 1);
  }
  SNAC_HIP(hipGetLastError());
}

void Net::loss_backward(hipStream_t s, int B, const float *t, float *loss_dev) {
  const long pl0 = (long)plane(0);
  hipLaunchKernelGGL(k_loss_sums, dim3(B), dim3(256), 0, s, pl0, y_, t, loss_part_);
  hipLaunchKernelGGL(k_loss_grad, dim3(grid_for((long)B * pl0 * 16)), dim3(256), 0, s, B, pl0, y_, t, loss_part_, gy_,
                     loss_dev);
  SNAC_HIP(hipMemsetAsync(g32_, 0, nparam_ * sizeof(float), s));
  const float *in[NET_LAYERS] = {act_[0], pool_[0], pool_[1], pool_[2], sum_[2], sum_[1], sum_[0]};
  // g of each layer's output: layer 6 gy_, layer l < 6 gact_[l + 1]
  for (int l = NET_LAYERS - 1; l >= 0; l--) {
    const int lv = NET_LEVEL[l], h = H_ >> lv, w = W_ >> lv;
    float *g = l == NET_LAYERS - 1 ? gy_ : gact_[l + 1];
    const long nout = (long)B * h * w * NET_COUT[l];
    if (l < NET_LAYERS - 1)
      hipLaunchKernelGGL(k_relu_bwd, dim3(grid_for(nout)), dim3(256), 0, s, nout, act_[l + 1], g);
    wgrad_any(s, NET_CIN[l], NET_COUT[l], B, h, w, in[l], g, g32_ + woff_[l]);
    bias_grad_any(s, NET_COUT[l], (long)B * h * w, g, g32_ + boff_[l]);
    if (l == 0) break;
    // gradient of this layer's input (flipped weights, no bias, no ReLU)
    float *gin = l >= 4 ? gsum_[6 - l] : gpool_[l - 1];
    conv_any(s, NET_COUT[l], NET_CIN[l], B, h, w, g, wf_ + woff_[l], nullptr, gin, false);
    const int ch = NET_CIN[l];
    if (l >= 4) {
      // input = up(act_[l]) + skip act_[q + 1]: the skip gets gin, the
      // coarse branch the sum over each pixel's four children
      const int q = 6 - l;
      const long n = (long)B * h * w * ch;
      hipLaunchKernelGGL(k_copy, dim3(grid_for(n)), dim3(256), 0, s, n, gin, gact_[q + 1]);
      const long nc = (long)B * (h / 2) * (w / 2) * ch;
      hipLaunchKernelGGL(k_up_bwd, dim3(grid_for(nc)), dim3(256), 0, s, B, h / 2, w / 2, ch, gin, gact_[l]);
    } else {
      // input = pool(act_[l]): every fine pixel gets a quarter
      const long n = (long)B * (2 * h) * (2 * w) * ch;
      hipLaunchKernelGGL(k_pool_bwd_add, dim3(grid_for(n)), dim3(256), 0, s, B, 2 * h, 2 * w, ch, gin, gact_[l]);
    }
  }
  SNAC_HIP(hipGetLastError());
}

void Net::adam_step(hipStream_t s, float lr) {
  step_++;
  const float c1 = 1.f / (1.f - std::pow(0.9f, (float)step_)), c2 = 1.f / (1.f - std::pow(0.999f, (float)step_));
  hipLaunchKernelGGL(k_adam, dim3(grid_for((long)nparam_)), dim3(256), 0, s, (long)nparam_, p32_, g32_, m_, v_, lr, c1,
                     c2);
  refresh_bf16(s);
}

}  // namespace snac
