// Kernel launchers of the learned predictor (NHWC, BF16 activations and
// weights, FP32 gradients and accumulation).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace piv {

// A 3x3 convolution with zero padding 1: input hi x wi, output ho x wo =
// hi/stride x wi/stride, co output channels, over `batch` images.
struct ConvGeom {
  int batch;
  int hi, wi;
  int ho, wo;
  int stride;
  int co;
};

enum ConvOut { kOutBf16 = 0, kOutF32 = 1, kOutF32Planar = 2 };

// Forward: input channel c < c0 comes from s0 (read at (y/2, x/2) when up0),
// c >= c0 from s1 (channel c - c0).  w: BF16 [co][9][c0+c1]; bias FP32 [co].
// out: NHWC BF16, NHWC FP32 or per-image planar FP32 [n][co][y][x]
// (ConvOut); ReLU optional.
hipError_t launch_conv_fwd(const ConvGeom& g, const uint16_t* s0, int c0, int up0, const uint16_t* s1, int c1,
                           const uint16_t* w, const float* bias, void* out, int out_mode, int relu, hipStream_t s);

// Input gradient dx (FP32 [batch][hi][wi][ci], written or accumulated) from
// dy (FP32 [batch][ho][wo][co]) and wt (BF16 [ci][9][co]).
hipError_t launch_conv_bwd_data(const ConvGeom& g, int ci, const float* dy, const uint16_t* wt, float* dx,
                                int accumulate, hipStream_t s);

// Weight and bias gradients (written) through a partial-sum workspace of
// weight_grad_part_floats(g, c0 + c1) floats.
size_t weight_grad_part_floats(const ConvGeom& g, int ci);
hipError_t launch_conv_bwd_weight(const ConvGeom& g, const uint16_t* s0, int c0, int up0, const uint16_t* s1, int c1,
                                  const float* dy, float* part, float* dw, float* db, hipStream_t s);

hipError_t launch_relu_mask(float* g, const uint16_t* act, long count, hipStream_t s);
hipError_t launch_split_grad(const float* t, int batch, int hi, int wi, int c0, int c1, float* d0, float* d1,
                             hipStream_t s);
size_t loss_part_floats(long count);
// out/target: per-image planar [batch][2][h][w] (count = batch*2*hw);
// weight: [batch][h][w] or null; grad: NHWC [batch][h][w][2].
hipError_t launch_loss(const float* out, const float* target, const float* weight, long count, long hw, float* grad,
                       float* part, float* loss, hipStream_t s);
hipError_t launch_adam(float* p, const float* g, float* m, float* v, long count, float lr, float b1, float b2,
                       float eps, int step, hipStream_t s);
hipError_t launch_weight_convert(const float* p, int co, int ci, uint16_t* w, uint16_t* wt, hipStream_t s);

// Network input from a frame pair: f x f box means, each frame standardised
// to zero mean and unit variance, BF16 [ht/f][wd/f][2].
size_t prepare_scratch_bytes(int ht, int wd, int f);
hipError_t launch_prepare_input(const void* fa, const void* fb, int dtype, int ht, int wd, int f, void* scratch,
                                uint16_t* out, hipStream_t s);

}  // namespace piv

struct piv_cnn;

namespace piv {

struct CnnInfo {
  int in_h, in_w;    // network input (frame / kCnnFactor)
  int out_h, out_w;  // output grid (input / 2): node (i, j) at frame (2f j + f, 2f i + f)
};
constexpr int kCnnFactor = 8;

CnnInfo cnn_info(const piv_cnn* net);
// Forward pass of `batch` inputs (BF16 [batch][in_h][in_w][2]); the output
// stays in the network's buffer, per-image planar FP32 [batch][2][out_h][out_w].
int cnn_enqueue_forward(piv_cnn* net, const uint16_t* input, int batch, hipStream_t s);
const float* cnn_output(const piv_cnn* net);

}  // namespace piv
