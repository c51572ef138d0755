/*
 * net_oracle.c -- CPU restatement of the learned preconditioner's network
 * (snac_amd/csrc/net.hpp: a 4-level convolutional encoder-decoder per
 * Fourier mode) and of its backward pass, for the gradient checks of the
 * GPU's BF16 MFMA kernels.
 *
 * TEST INFRASTRUCTURE ONLY (the checker; never shipped).  The network has no
 * counterpart in the reference.  Pinned by central finite differences of the
 * loss (tests/test_oracle_net.py) with exact operands (round = 0); with
 * round = 1 the convolution operands are rounded to BF16 as the MFMA
 * kernels round them (accumulation here is FP64, there FP32).
 *
 * Layout: NHWC, FP32 storage; parameters flattened per layer: weights
 * [tap][cout][cin] then biases [cout]; channel counts padded to 16.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define NL 7
static const int CIN[NL] = {16, 16, 32, 64, 64, 32, 16};
static const int COUT[NL] = {16, 32, 64, 64, 32, 16, 16};
static const int LEV[NL] = {0, 1, 2, 3, 2, 1, 0};

static float bf16r(float x, int round) {
  if (!round) return x;
  unsigned u;
  memcpy(&u, &x, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) u = (u >> 16 | ((u & 0xffffu) ? 0x40u : 0u)) << 16;
  else { u += 0x7fffu + ((u >> 16) & 1u); u &= 0xffff0000u; }
  float f;
  memcpy(&f, &u, 4);
  return f;
}

long on_num_params(void) {
  long n = 0;
  for (int l = 0; l < NL; l++) n += 9L * CIN[l] * COUT[l] + COUT[l];
  return n;
}

static void offsets(long *wo, long *bo) {
  long off = 0;
  for (int l = 0; l < NL; l++) { wo[l] = off; off += 9L * CIN[l] * COUT[l]; bo[l] = off; off += COUT[l]; }
}

/* out = act(b + conv(in, w)); zero padding */
static void conv(int B, int H, int W, int ci, int co, const float *in, const float *w, const float *b, float *out,
                 int relu, int round) {
  for (int bi = 0; bi < B; bi++)
    for (int y = 0; y < H; y++)
      for (int x = 0; x < W; x++)
        for (int o = 0; o < co; o++) {
          double s = 0.0;
          for (int tap = 0; tap < 9; tap++) {
            int yy = y + tap / 3 - 1, xx = x + tap % 3 - 1;
            if (yy < 0 || yy >= H || xx < 0 || xx >= W) continue;
            const float *src = in + (((long)bi * H + yy) * W + xx) * ci;
            const float *wp = w + ((long)tap * co + o) * ci;
            for (int c = 0; c < ci; c++) s += (double)bf16r(src[c], round) * (double)bf16r(wp[c], round);
          }
          float v = (float)s + (b ? b[o] : 0.f);
          if (relu && v < 0.f) v = 0.f;
          out[(((long)bi * H + y) * W + x) * co + o] = v;
        }
}

static void pool(int B, int Ho, int Wo, int C, const float *in, float *out) {
  for (int bi = 0; bi < B; bi++)
    for (int y = 0; y < Ho; y++)
      for (int x = 0; x < Wo; x++)
        for (int c = 0; c < C; c++) {
          long Wi = 2L * Wo, base = (((long)bi * 2 * Ho + 2 * y) * Wi + 2 * x) * C + c;
          out[(((long)bi * Ho + y) * Wo + x) * C + c] = 0.25f * (in[base] + in[base + C] + in[base + Wi * C] + in[base + Wi * C + C]);
        }
}

static void up_add(int B, int H, int W, int C, const float *coarse, const float *skip, float *out) {
  for (int bi = 0; bi < B; bi++)
    for (int y = 0; y < H; y++)
      for (int x = 0; x < W; x++)
        for (int c = 0; c < C; c++) {
          long q = (((long)bi * H + y) * W + x) * C + c;
          out[q] = coarse[(((long)bi * (H / 2) + y / 2) * (W / 2) + x / 2) * C + c] + skip[q];
        }
}

typedef struct {
  float *act[NL + 1], *pin[NL], *y;
} fwd_t;

static float *zalloc(long n) { return calloc((size_t)(n > 0 ? n : 1), sizeof(float)); }

static void forward(int B, int H, int W, const float *p, const float *x16, fwd_t *f, int round) {
  long wo[NL], bo[NL];
  offsets(wo, bo);
  f->act[0] = zalloc((long)B * H * W * 16);
  memcpy(f->act[0], x16, sizeof(float) * (size_t)B * H * W * 16);
  for (int l = 0; l < NL; l++) {
    int h = H >> LEV[l], w = W >> LEV[l];
    float *in;
    if (l == 0) in = f->act[0];
    else if (l <= 3) { in = zalloc((long)B * h * w * CIN[l]); pool(B, h, w, CIN[l], f->act[l], in); }
    else { in = zalloc((long)B * h * w * CIN[l]); up_add(B, h, w, CIN[l], f->act[l], f->act[6 - l + 1], in); }
    f->pin[l] = in;
    float *out = zalloc((long)B * h * w * COUT[l]);
    conv(B, h, w, CIN[l], COUT[l], in, p + wo[l], p + bo[l], out, l < NL - 1, round);
    if (l < NL - 1) f->act[l + 1] = out; else f->y = out;
  }
}

static void free_fwd(fwd_t *f) {
  for (int l = 0; l <= NL; l++) free(f->act[l]);
  for (int l = 1; l < NL; l++) free(f->pin[l]);
  free(f->y);
}

/* y (channel 0 of B images of 16 channels): x2 is B planes of 2 channels */
static float *expand(int B, int H, int W, const float *x2) {
  float *x16 = zalloc((long)B * H * W * 16);
  for (long p = 0; p < (long)B * H * W; p++) { x16[p * 16] = x2[2 * p]; x16[p * 16 + 1] = x2[2 * p + 1]; }
  return x16;
}

int on_forward(int B, int H, int W, const float *p, const float *x2, float *y, int round) {
  if (H % 8 || W % 8) return -1;
  float *x16 = expand(B, H, W, x2);
  fwd_t f;
  memset(&f, 0, sizeof f);
  forward(B, H, W, p, x16, &f, round);
  for (long q = 0; q < (long)B * H * W; q++) y[q] = f.y[q * 16];
  free_fwd(&f);
  free(x16);
  return 0;
}

/* loss = mean_b |y_b - t_b|^2 / |t_b|^2 and its gradient w.r.t. p */
int on_loss_grad(int B, int H, int W, const float *p, const float *x2, const float *t, float *loss, float *grad,
                 int round) {
  if (H % 8 || W % 8) return -1;
  long wo[NL], bo[NL];
  offsets(wo, bo);
  long pl = (long)H * W;
  float *x16 = expand(B, H, W, x2);
  fwd_t f;
  memset(&f, 0, sizeof f);
  forward(B, H, W, p, x16, &f, round);
  float *gy = zalloc((long)B * pl * 16);
  double L = 0.0;
  for (int b = 0; b < B; b++) {
    double s1 = 0, s2 = 0;
    for (long q = 0; q < pl; q++) {
      double d = f.y[(b * pl + q) * 16] - t[b * pl + q];
      s1 += d * d;
      s2 += (double)t[b * pl + q] * t[b * pl + q];
    }
    if (s2 <= 0) s2 = 1;
    L += s1 / s2;
    for (long q = 0; q < pl; q++) gy[(b * pl + q) * 16] = (float)(2.0 * (f.y[(b * pl + q) * 16] - t[b * pl + q]) / (B * s2));
  }
  *loss = (float)(L / B);
  memset(grad, 0, sizeof(float) * (size_t)on_num_params());
  /* gradients of the activations act[1..6] (allocated lazily) */
  float *gact[NL + 1] = {0};
  for (int l = NL - 1; l >= 0; l--) {
    int h = H >> LEV[l], w = W >> LEV[l];
    long npx = (long)B * h * w;
    float *g = l == NL - 1 ? gy : gact[l + 1];
    if (l < NL - 1)
      for (long q = 0; q < npx * COUT[l]; q++) if (!(f.act[l + 1][q] > 0.f)) g[q] = 0.f;
    /* weight and bias gradients */
    const float *in = f.pin[l];
    for (int tap = 0; tap < 9; tap++)
      for (int o = 0; o < COUT[l]; o++)
        for (int c = 0; c < CIN[l]; c++) {
          double s = 0;
          for (int bi = 0; bi < B; bi++)
            for (int y = 0; y < h; y++)
              for (int x = 0; x < w; x++) {
                int yy = y + tap / 3 - 1, xx = x + tap % 3 - 1;
                if (yy < 0 || yy >= h || xx < 0 || xx >= w) continue;
                s += (double)bf16r(g[(((long)bi * h + y) * w + x) * COUT[l] + o], round) *
                     (double)bf16r(in[(((long)bi * h + yy) * w + xx) * CIN[l] + c], round);
              }
          grad[wo[l] + ((long)tap * COUT[l] + o) * CIN[l] + c] = (float)s;
        }
    for (int o = 0; o < COUT[l]; o++) {
      double s = 0;
      for (long q = 0; q < npx; q++) s += g[q * COUT[l] + o];
      grad[bo[l] + o] = (float)s;
    }
    if (l == 0) break;
    /* gradient of the layer input: sum over taps of g at the shifted pixel */
    float *gin = zalloc(npx * CIN[l]);
    for (int bi = 0; bi < B; bi++)
      for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++)
          for (int c = 0; c < CIN[l]; c++) {
            double s = 0;
            for (int tap = 0; tap < 9; tap++) {
              int yy = y - (tap / 3 - 1), xx = x - (tap % 3 - 1);
              if (yy < 0 || yy >= h || xx < 0 || xx >= w) continue;
              for (int o = 0; o < COUT[l]; o++)
                s += (double)bf16r(g[(((long)bi * h + yy) * w + xx) * COUT[l] + o], round) *
                     (double)bf16r(p[wo[l] + ((long)tap * COUT[l] + o) * CIN[l] + c], round);
            }
            gin[(((long)bi * h + y) * w + x) * CIN[l] + c] = (float)s;
          }
    int ch = CIN[l];
    if (l >= 4) {
      int q = 6 - l;
      /* skip branch act[q+1]; coarse branch act[l] */
      long nfine = npx * ch;
      if (!gact[q + 1]) gact[q + 1] = zalloc(nfine);
      for (long r = 0; r < nfine; r++) gact[q + 1][r] += gin[r];
      int hc = h / 2, wc = w / 2;
      gact[l] = zalloc((long)B * hc * wc * ch);
      for (int bi = 0; bi < B; bi++)
        for (int y = 0; y < h; y++)
          for (int x = 0; x < w; x++)
            for (int c = 0; c < ch; c++)
              gact[l][(((long)bi * hc + y / 2) * wc + x / 2) * ch + c] += gin[(((long)bi * h + y) * w + x) * ch + c];
    } else {
      int hf = 2 * h, wf = 2 * w;
      if (!gact[l]) gact[l] = zalloc((long)B * hf * wf * ch);
      for (int bi = 0; bi < B; bi++)
        for (int y = 0; y < hf; y++)
          for (int x = 0; x < wf; x++)
            for (int c = 0; c < ch; c++)
              gact[l][(((long)bi * hf + y) * wf + x) * ch + c] += 0.25f * gin[(((long)bi * h + y / 2) * w + x / 2) * ch + c];
    }
    free(gin);
  }
  for (int l = 0; l <= NL; l++) free(gact[l]);
  free(gy);
  free_fwd(&f);
  free(x16);
  return 0;
}
