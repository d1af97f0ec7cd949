// Convolutional network of the learned Schur-complement correction: four 3x3
// convolution layers (4 -> 32 -> 32 -> 32 -> 1 channels, ReLU between) on the
// pressure-node grid, bfloat16 weights and activations, FP32 accumulation.
// Forward and backward passes and the Adam update are hand-written kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mw_solver.h"

namespace mw {

constexpr int CNN_HID = 32;
constexpr int CNN_CIN = 4;
// packed FP32 parameters: W1 [9][32][4], b1, W2 [9][32][32], b2, W3, b3, W4 [9][1][32], b4
constexpr int CNN_OFF_W1 = 0, CNN_OFF_W2 = 1184, CNN_OFF_W3 = 10432, CNN_OFF_W4 = 19680;
constexpr int CNN_NPARAM = 19969;
// bfloat16 weight copies: W1, W2, W3, W4 in the FP32 layout, then W2 and W3
// transposed and tap-flipped for the input-gradient convolutions
constexpr int CNN_BF_W1 = 0, CNN_BF_W2 = 1152, CNN_BF_W3 = 10368, CNN_BF_W4 = 19584,
              CNN_BF_W2T = 19872, CNN_BF_W3T = 29088, CNN_NBF = 38304;
constexpr int CNN_PART_BLOCKS = 256;  // blocks of the weight-gradient reductions

struct Cnn {
  int H = 0, W = 0, maxB = 0;
  long npix = 0;
  long t_adam = 0;
  DBuf<float> params, grads, m, v, part;
  DBuf<uint16_t> wbf;
  DBuf<float> X0;                       // maxB * npix * 4 (kept for the backward pass)
  DBuf<uint16_t> A1, A2, A3, dZa, dZb;  // maxB * npix * 32
  int init(int H, int W, int maxB);
  int set_params(const float* host, hipStream_t st);
  int get(const DBuf<float>& b, float* host, hipStream_t st);
  void refresh_bf16(hipStream_t st);
  // C (B*npix) from X0 (B*npix*4, NHWC); X0 may be this->X0
  int forward(int B, const float* x0, float* C, hipStream_t st);
  // grads = d(sum dC*C)/dparams of the last forward pass
  int backward(int B, const float* dC, hipStream_t st);
  void adam(double lr, double beta1, double beta2, double eps, hipStream_t st);
};

}  // namespace mw
