// Learned right preconditioner for flexible GMRES: a 3-level convolutional
// U-Net (32 base channels) on a coarse cell grid, BF16 weights and
// activations, FP32 accumulation on MFMA (v_mfma_f32_32x32x16_bf16).  Forward,
// backward and the Adam update all run on the GPU.
//
// Architecture (channels-last tensors [sample][z][y][x][channel], G^3 grid):
//   in(16: r_hat[8], tau_hat, kappa_hat, 0...) -> enc0a(32) -> enc0b(32)=skip0
//   avgpool2 -> enc1a(64) -> enc1b(64)=skip1 -> avgpool2 -> bota(128) -> botb(128)
//   upsample2 ++ skip1 (192) -> dec1a(64) -> dec1b(64)
//   upsample2 ++ skip0 (96)  -> dec0a(32) -> dec0b(32) -> head 1x1 (8, FP32)
// Every 3x3x3 convolution has a bias and a ReLU; zero padding at the grid edge.
//
// Per coarse cell the features (FP32, kUnetFeatures values) are r_hat[8]
// (the residual restricted to the cell's eight octants, divided by its RMS;
// octant k = kx + 2 ky + 4 kz), tau_hat, kappa_hat (normalised optical
// thickness and log absorption) and g (a fixed amplification prior).  The
// eight outputs are the correction at the cell's eight corners.  The network
// is evaluated on +r_hat and -r_hat and the prediction is the odd part
// (N(+r) - N(-r)) / 2, optionally plus a skip connection r_hat and scaled by
// the prior g (set_mode), so the preconditioner is odd and, with the RMS
// scaling, positively homogeneous in the residual.
//
// Parameter vector (FP32 master copy; BF16 mirror for the MFMA operands):
//   per convolution: W[27 taps][cout][cin] then bias[cout];  tap = 9(dz+1)+3(dy+1)+(dx+1)
//   head: w[8][32] then b[8].
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "ou_common.hpp"

namespace ou {

constexpr int kUnetOut = 8;        // corner values per cell
constexpr int kUnetFeatures = 12;  // r_hat[8], tau_hat, kappa_hat, g, (unused)
constexpr int kUnetFeatTau = 8, kUnetFeatKappa = 9, kUnetFeatGain = 10;
constexpr int kUnetInPad = 16;     // network input channels (10 used, zero padded)
constexpr int kUnetInReal = 10;
constexpr int kUnetHead = kUnetOut * 32 + kUnetOut;  // head weights and biases
constexpr int kUnetConvs = 10;

struct UnetConv {
  int cin, cout;
  long w_off, b_off;
};

class UNet {
 public:
  UNet(int grid, int max_batch, uint64_t seed);
  ~UNet();
  UNet(const UNet &) = delete;
  UNet &operator=(const UNet &) = delete;

  int grid() const { return G_; }
  int max_batch() const { return max_batch_; }
  long n_params() const { return n_params_; }
  const UnetConv &conv(int i) const { return conv_[i]; }

  void get_params(float *h) const;
  void set_params(const float *h);

  // pred[b][cell][k] = gain (skip r_hat[k] + (N(+r_hat) - N(-r_hat))[k] / 2), k < 8,
  // for B samples of device features feat[b][cell][kUnetFeatures] (FP32).
  void predict(const float *d_feat, int B, float *d_pred, hipStream_t s);
  // Loss sum_b |pred_b - t_b|^2 / |t_b|^2 / B (targets [b][cell][8]) and its
  // gradient into d_grad().
  void loss_and_grad(const float *d_feat, const float *d_target, int B, hipStream_t s);
  // Adam (beta1 0.9, beta2 0.999, eps 1e-8) on the FP32 master parameters,
  // then the BF16 mirror is refreshed.
  void adam_step(float lr, hipStream_t s);
  const float *d_grad() const { return grad_; }
  const double *d_loss() const { return loss_; }
  int adam_t() const { return adam_t_; }
  // Prediction form: bit 0 adds the skip r_hat, bit 1 scales by the prior g
  // (both off: pred = (N(+r) - N(-r)) / 2).
  void set_mode(int mode) {
    if (mode < 0 || mode > 3) fail(OU_ERR_ARG, "U-Net mode must be in [0, 3]");
    mode_ = mode;
  }
  int mode() const { return mode_; }

 private:
  void forward(const float *d_feat, int B, hipStream_t s);
  void conv_forward(int layer, const uint16_t *x, int ldx, int xoff, uint16_t *y, int ldy,
                    int yoff, int D, int nb, hipStream_t s);
  void conv_backward(int layer, const uint16_t *x, int ldx, int xoff, const uint16_t *dy,
                     int ldy, int yoff, uint16_t *dx, int lddx, int dxoff, const uint16_t *mask,
                     int ldm, int moff, int D, int nb, hipStream_t s);
  void refresh_bf16(hipStream_t s);

  int G_, max_batch_;
  long n_params_ = 0;
  UnetConv conv_[kUnetConvs];
  long head_w_off_ = 0, head_b_off_ = 0;
  int adam_t_ = 0;
  int mode_ = 0;

  // device memory
  float *param_ = nullptr, *grad_ = nullptr, *adam_m_ = nullptr, *adam_v_ = nullptr;
  uint16_t *wbf_ = nullptr;    // BF16 mirror of param_
  uint16_t *wflip_ = nullptr;  // one layer's flipped weights for the data gradient
  float *partial_ = nullptr;   // weight/bias gradient partial sums
  long partial_size_ = 0;
  double *loss_ = nullptr;     // [1 + 2 max_batch]: loss, then per-sample num/den
  float *loss_part_ = nullptr;
  // activations and their gradients (BF16), 2*max_batch samples
  uint16_t *in0_ = nullptr, *e0a_ = nullptr, *cat0_ = nullptr, *p0_ = nullptr, *e1a_ = nullptr,
           *cat1_ = nullptr, *p1_ = nullptr, *ba_ = nullptr, *bb_ = nullptr, *d1a_ = nullptr,
           *d1b_ = nullptr, *d0a_ = nullptr, *d0b_ = nullptr;
  uint16_t *g_e0a_ = nullptr, *g_cat0_ = nullptr, *g_p0_ = nullptr, *g_e1a_ = nullptr,
           *g_cat1_ = nullptr, *g_p1_ = nullptr, *g_ba_ = nullptr, *g_bb_ = nullptr,
           *g_d1a_ = nullptr, *g_d1b_ = nullptr, *g_d0a_ = nullptr, *g_d0b_ = nullptr;
  float *out_ = nullptr, *g_out_ = nullptr;
  std::vector<void *> allocs_;
  void *alloc(size_t bytes);
};

}  // namespace ou
