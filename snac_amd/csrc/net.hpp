// The learned preconditioner's network: a per-mode convolutional
// encoder-decoder (4 levels, 16-64 channels) on the 2D plane of one Fourier
// mode, with BF16 MFMA convolutions accumulating in FP32 and its own
// backward pass and Adam optimiser.  No counterpart in the reference (its
// solver is HYPRE's MG-PCG); the network only ever preconditions a flexible
// CG that the plain FP64 residual judges.
//
// Architecture (H, W multiples of 8; NHWC FP32 activations; 3x3 convs with
// zero padding; every hidden conv followed by ReLU):
//   x  (H,   16: r/|r|, lambda feature, 14 zero channels)
//   e1 = conv0(x)                 16 ch   H
//   e2 = conv1(pool(e1))          32 ch   H/2
//   e3 = conv2(pool(e2))          64 ch   H/4
//   e4 = conv3(pool(e3))          64 ch   H/8
//   d3 = conv4(up(e4) + e3)       32 ch   H/4
//   d2 = conv5(up(d3) + e2)       16 ch   H/2
//   y  = conv6(up(d2) + e1)       channel 0 of 16, linear   H
// pool = 2x2 average, up = nearest.  Weights are kept in FP32 (the master
// copy the optimiser updates) and in BF16 in the layout the MFMA operands
// read ([tap][cout][cin], and the flipped [tap][cin][cout] for the data
// gradient).
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

namespace snac {

constexpr int NET_LAYERS = 7;
// input / output channels (padded to 16) and level (0 = full resolution)
constexpr int NET_CIN[NET_LAYERS] = {16, 16, 32, 64, 64, 32, 16};
constexpr int NET_COUT[NET_LAYERS] = {16, 32, 64, 64, 32, 16, 16};
constexpr int NET_LEVEL[NET_LAYERS] = {0, 1, 2, 3, 2, 1, 0};

class Net {
 public:
  Net(int H, int W, int maxbatch, unsigned seed);
  ~Net();
  Net(const Net &) = delete;
  Net &operator=(const Net &) = delete;
  int H() const { return H_; }
  int W() const { return W_; }
  int maxbatch() const { return mb_; }
  // parameters, flattened: for each layer its weights [tap][cout][cin]
  // (FP32) then its biases [cout]
  size_t num_params() const { return nparam_; }
  void get_params(float *host) const;
  void set_params(const float *host);
  // the input planes (B images of 16 channels, NHWC, device) the caller
  // fills before forward; the output (channel 0 of B images of 16
  // channels) after it
  float *input() { return act_[0]; }
  const float *output() const { return y_; }
  // input from B planes of 2 channels ([b][y][x][2], device); output channel
  // 0 into B planes (device)
  void set_input2(hipStream_t s, int B, const float *x2);
  void get_output(hipStream_t s, int B, float *y) const;
  void forward(hipStream_t s, int B);
  // loss = mean over images of |y - t|^2 / |t|^2 (t: B planes H x W,
  // device); gradients of every parameter into the gradient buffer; the
  // loss is returned through loss_dev (one float, device)
  void loss_backward(hipStream_t s, int B, const float *t, float *loss_dev);
  void grads_to_host(float *host) const;
  void adam_step(hipStream_t s, float lr);

 private:
  int H_, W_, mb_;
  size_t nparam_ = 0;
  size_t woff_[NET_LAYERS], boff_[NET_LAYERS];
  float *p32_ = nullptr, *g32_ = nullptr, *m_ = nullptr, *v_ = nullptr;  // params, grads, Adam moments
  short *wb_ = nullptr, *wf_ = nullptr;                                  // BF16 weights, flipped
  int step_ = 0;
  // activations: act_[0] = x; per layer the ReLU output; pooled inputs;
  // up+skip sums; the output
  float *act_[NET_LAYERS + 1] = {};
  float *pool_[3] = {}, *sum_[3] = {}, *y_ = nullptr;
  // gradients of the activations
  float *gact_[NET_LAYERS + 1] = {}, *gpool_[3] = {}, *gsum_[3] = {}, *gy_ = nullptr;
  float *loss_part_ = nullptr;
  std::vector<void *> owned_;
  float *alloc(size_t n);
  size_t plane(int level) const { return (size_t)(H_ >> level) * (W_ >> level); }
  void refresh_bf16(hipStream_t s);
};

// bf16 of an FP32 value, rounded to nearest even (the MFMA operands)
float bf16_round(float x);
// convolution kernel: 0 per-tap global reads, 1 LDS-staged input (default)
void net_set_conv_variant(int v);

// ---- the MG-PCG planes <-> network planes (kernels_net.hip) ----
// Group g of a plan vector is the na x ns plane (row a, column s, s
// contiguous) of mode g; a network plane is nH x nW (>= na, ns, multiples
// of 8), zero outside.  Each plane is scaled by its RMS, sqrt(ss[2g] /
// nact), so the network sees unit-RMS inputs (ss: per-group sums of squares,
// stride 2, as the CG's dot products store them); channel 1 carries the
// mode's eigenvalue feature lam[g] / lam_ref.
// x16: the network input (ng planes of 16 channels)
void launch_net_in(hipStream_t s, int ng, int na, int ns, int nH, int nW, const float *r, const double *ss,
                   double nact, const double *lam, double lam_ref, float *x16);
// z = scale_g * y (channel 0 of the network output) on the active cells,
// 0 elsewhere
void launch_net_out(hipStream_t s, int ng, int na, int ns, int nH, int nW, const float *y16, const double *ss,
                    double nact, const unsigned char *act, float *z);
// training pairs from a solve: the input half from the right-hand side
// (FP64 plan vector b with sums of squares bb), the target half x / scale
// from the solution
void launch_net_pair_in(hipStream_t s, int ng, int na, int ns, int nH, int nW, const double *b, const double *bb,
                        double nact, const double *lam, double lam_ref, float *x2);
void launch_net_pair_out(hipStream_t s, int ng, int na, int ns, int nH, int nW, const double *x, const double *bb,
                         double nact, float *t);
// stall |= some group not done whose r.r fell by less than ratio2 since
// the last call; then prev = rr (both stride 2)
void launch_net_stall(hipStream_t s, int ng, const double *rr, double *prev, const int *done, double ratio2,
                      int *stall);

}  // namespace snac
