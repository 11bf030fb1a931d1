// The coarse-displacement network: buffers, forward, backward, Adam, and
// its C ABI.  No counterpart in the reference (its predictor for window
// deformation is always the previous pass, gpu_process.py:620-625).
//
// Architecture (input R_h x R_w x 2 = box-averaged, standardised frames A, B
// at 1/8 of the frame; 3x3 convolutions, ReLU after all but the head):
//   c1  2 -> 32   R        c5  64 -> 64   R/4 (skip)    d2 64 -> 64  R/4
//   c2 32 -> 48   R/2 (s2) c6  64 -> 128  R/8 (s2)      d3 up(d2) ++ c3 = 112 -> 48  R/2
//   c3 48 -> 48   R/2 (skip) c7 128 -> 128 R/8          head 48 -> 2  R/2 (FP32, planar u, v)
//   c4 48 -> 64   R/4 (s2) d1 up(c7) ++ c5 = 192 -> 64  R/4
// 518,226 parameters; output node (i, j) sits at frame (16 j + 8, 16 i + 8).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <vector>

#include "../../include/openpiv_hip.h"
#include "piv_cnn.h"
#include "piv_kernels.h"

namespace piv {
int capi_fail(int code, const char* fmt, ...);
}

namespace {

struct LayerSpec {
  int s0, s1, up0, co, stride, relu;  // source buffers (-1 none), x2 upsampling of s0
};
// buffer 0 = network input, buffer L+1 = output of layer L
constexpr LayerSpec kArch[] = {
    {0, -1, 0, 32, 1, 1},  {1, -1, 0, 48, 2, 1},  {2, -1, 0, 48, 1, 1},  {3, -1, 0, 64, 2, 1},
    {4, -1, 0, 64, 1, 1},  {5, -1, 0, 128, 2, 1}, {6, -1, 0, 128, 1, 1}, {7, 5, 1, 64, 1, 1},
    {8, -1, 0, 64, 1, 1},  {9, 3, 1, 48, 1, 1},   {10, -1, 0, 2, 1, 0},
};
constexpr int kLayers = sizeof(kArch) / sizeof(kArch[0]);
constexpr int kBuffers = kLayers + 1;

struct Layer {
  LayerSpec spec;
  int c0, c1;
  piv::ConvGeom g;  // batch filled per call
  size_t w_off, b_off;
};

}  // namespace

struct piv_cnn {
  int device = 0;
  int in_h = 0, in_w = 0, max_batch = 0;
  Layer L[kLayers];
  int buf_h[kBuffers], buf_w[kBuffers], buf_c[kBuffers];
  size_t n_params = 0;
  float* params = nullptr;
  float* grads = nullptr;
  float* adam_m = nullptr;
  float* adam_v = nullptr;
  uint16_t* wbf = nullptr;   // BF16 weights [co][9][ci], at w_off
  uint16_t* wtbf = nullptr;  // BF16 transposed [ci][9][co], at w_off
  uint16_t* act[kBuffers] = {};  // act[0] unused (caller's input)
  float* out = nullptr;          // head output, planar FP32
  float* gact[kBuffers] = {};    // gradients of buffers 1..kLayers (NHWC FP32)
  float* concat_tmp = nullptr;
  float* part = nullptr;
  float* loss_part = nullptr;
  float* loss = nullptr;
  void* prep_scratch = nullptr;
  int step = 0;
};

namespace {

using piv::capi_fail;

#define CNN_TRY(expr)                                                                            \
  do {                                                                                           \
    hipError_t _e = (expr);                                                                      \
    if (_e != hipSuccess) return capi_fail(PIV_E_HIP, "%s: %s", #expr, hipGetErrorString(_e)); \
  } while (0)

piv::ConvGeom geom(const piv_cnn* n, int l, int batch) {
  piv::ConvGeom g = n->L[l].g;
  g.batch = batch;
  return g;
}

const uint16_t* src_ptr(const piv_cnn* n, int buf, const uint16_t* input) {
  if (buf < 0) return nullptr;
  return buf == 0 ? input : n->act[buf];
}

int forward(piv_cnn* n, const uint16_t* input, int batch, hipStream_t s) {
  for (int l = 0; l < kLayers; ++l) {
    const Layer& L = n->L[l];
    const bool head = l == kLayers - 1;
    void* out = head ? (void*)n->out : (void*)n->act[l + 1];
    CNN_TRY(piv::launch_conv_fwd(geom(n, l, batch), src_ptr(n, L.spec.s0, input), L.c0, L.spec.up0,
                                 src_ptr(n, L.spec.s1, input), L.c1, n->wbf + L.w_off, n->params + L.b_off, out,
                                 head ? piv::kOutF32Planar : piv::kOutBf16, L.spec.relu, s));
  }
  return PIV_OK;
}

int backward(piv_cnn* n, const uint16_t* input, int batch, hipStream_t s) {
  bool written[kBuffers] = {};
  written[kBuffers - 1] = true;  // head output gradient from the loss
  for (int l = kLayers - 1; l >= 0; --l) {
    const Layer& L = n->L[l];
    const piv::ConvGeom g = geom(n, l, batch);
    float* dy = n->gact[l + 1];
    if (L.spec.relu)
      CNN_TRY(piv::launch_relu_mask(dy, n->act[l + 1], (long)batch * g.ho * g.wo * g.co, s));
    CNN_TRY(piv::launch_conv_bwd_weight(g, src_ptr(n, L.spec.s0, input), L.c0, L.spec.up0,
                                        src_ptr(n, L.spec.s1, input), L.c1, dy, n->part, n->grads + L.w_off,
                                        n->grads + L.b_off, s));
    if (L.spec.s0 == 0) continue;  // no gradient for the network input
    const int ci = L.c0 + L.c1;
    if (L.spec.s1 < 0 && !L.spec.up0) {
      const int b = L.spec.s0;
      CNN_TRY(piv::launch_conv_bwd_data(g, ci, dy, n->wtbf + L.w_off, n->gact[b], written[b] ? 1 : 0, s));
      written[b] = true;
    } else {
      // the split writes both sources' gradients: it must be their first writer
      if (written[L.spec.s0] || (L.spec.s1 >= 0 && written[L.spec.s1]))
        return capi_fail(PIV_E_INVALID, "cnn: concatenated layer %d is not the first consumer in backward", l);
      CNN_TRY(piv::launch_conv_bwd_data(g, ci, dy, n->wtbf + L.w_off, n->concat_tmp, 0, s));
      CNN_TRY(piv::launch_split_grad(n->concat_tmp, batch, g.hi, g.wi, L.c0, L.c1, n->gact[L.spec.s0],
                                     L.spec.s1 >= 0 ? n->gact[L.spec.s1] : nullptr, s));
      written[L.spec.s0] = true;
      if (L.spec.s1 >= 0) written[L.spec.s1] = true;
    }
  }
  return PIV_OK;
}

int convert_weights(piv_cnn* n, hipStream_t s) {
  for (int l = 0; l < kLayers; ++l) {
    const Layer& L = n->L[l];
    CNN_TRY(piv::launch_weight_convert(n->params + L.w_off, L.spec.co, L.c0 + L.c1, n->wbf + L.w_off,
                                       n->wtbf + L.w_off, s));
  }
  return PIV_OK;
}

void release(piv_cnn* n) {
  if (!n) return;
  (void)hipSetDevice(n->device);
  for (void* p : {(void*)n->params, (void*)n->grads, (void*)n->adam_m, (void*)n->adam_v, (void*)n->wbf,
                  (void*)n->wtbf, (void*)n->out, (void*)n->concat_tmp, (void*)n->part, (void*)n->loss_part,
                  (void*)n->loss, n->prep_scratch})
    if (p) (void)hipFree(p);
  for (int b = 0; b < kBuffers; ++b) {
    if (n->act[b]) (void)hipFree(n->act[b]);
    if (n->gact[b]) (void)hipFree(n->gact[b]);
  }
  delete n;
}

int check_batch(const piv_cnn* n, int batch) {
  if (!n) return capi_fail(PIV_E_INVALID, "cnn: null network");
  if (batch < 1 || batch > n->max_batch) return capi_fail(PIV_E_INVALID, "cnn: batch %d outside 1..%d", batch, n->max_batch);
  return PIV_OK;
}

}  // namespace

namespace piv {

CnnInfo cnn_info(const piv_cnn* n) {
  return CnnInfo{n->in_h, n->in_w, n->buf_h[kBuffers - 1], n->buf_w[kBuffers - 1]};
}

int cnn_enqueue_forward(piv_cnn* n, const uint16_t* input, int batch, hipStream_t s) {
  int r = check_batch(n, batch);
  if (r) return r;
  return forward(n, input, batch, s);
}

const float* cnn_output(const piv_cnn* n) { return n->out; }

}  // namespace piv

extern "C" {

int piv_cnn_create(int in_height, int in_width, int max_batch, piv_cnn** out) {
  if (!out) return capi_fail(PIV_E_INVALID, "cnn: null output");
  *out = nullptr;
  if (in_height < 8 || in_width < 8 || in_height % 8 || in_width % 8)
    return capi_fail(PIV_E_INVALID, "cnn: input %d x %d must be a multiple of 8 in each direction", in_height,
                     in_width);
  if (max_batch < 1) return capi_fail(PIV_E_INVALID, "cnn: max_batch < 1");
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return capi_fail(PIV_E_NO_DEVICE, "cnn: no HIP device");
  piv_cnn* n = new piv_cnn();
  n->device = dev;
  n->in_h = in_height;
  n->in_w = in_width;
  n->max_batch = max_batch;
  n->buf_h[0] = in_height;
  n->buf_w[0] = in_width;
  n->buf_c[0] = 2;
  size_t off = 0, max_concat = 0, max_part = 0;
  for (int l = 0; l < kLayers; ++l) {
    Layer& L = n->L[l];
    L.spec = kArch[l];
    L.c0 = n->buf_c[L.spec.s0];
    L.c1 = L.spec.s1 >= 0 ? n->buf_c[L.spec.s1] : 0;
    const int hi = n->buf_h[L.spec.s0] * (L.spec.up0 ? 2 : 1), wi = n->buf_w[L.spec.s0] * (L.spec.up0 ? 2 : 1);
    if (L.spec.s1 >= 0 && (n->buf_h[L.spec.s1] != hi || n->buf_w[L.spec.s1] != wi)) {
      release(n);
      return capi_fail(PIV_E_INVALID, "cnn: skip resolution mismatch at layer %d", l);
    }
    L.g = piv::ConvGeom{max_batch, hi, wi, hi / L.spec.stride, wi / L.spec.stride, L.spec.stride, L.spec.co};
    n->buf_h[l + 1] = L.g.ho;
    n->buf_w[l + 1] = L.g.wo;
    n->buf_c[l + 1] = L.spec.co;
    const int ci = L.c0 + L.c1;
    L.w_off = off;
    off += (size_t)L.spec.co * 9 * ci;
    L.b_off = off;
    off += L.spec.co;
    off = (off + 3) / 4 * 4;  // keep each weight block 8-byte aligned in BF16
    if (L.spec.s1 >= 0 || L.spec.up0) max_concat = std::max(max_concat, (size_t)max_batch * hi * wi * ci);
    max_part = std::max(max_part, piv::weight_grad_part_floats(L.g, ci));
  }
  n->n_params = off;
  const long out_elems = (long)max_batch * n->buf_h[kBuffers - 1] * n->buf_w[kBuffers - 1] * 2;
  bool ok = hipMalloc((void**)&n->params, off * 4) == hipSuccess && hipMalloc((void**)&n->grads, off * 4) == hipSuccess &&
            hipMalloc((void**)&n->adam_m, off * 4) == hipSuccess && hipMalloc((void**)&n->adam_v, off * 4) == hipSuccess &&
            hipMalloc((void**)&n->wbf, off * 2) == hipSuccess && hipMalloc((void**)&n->wtbf, off * 2) == hipSuccess &&
            hipMalloc((void**)&n->out, out_elems * 4) == hipSuccess &&
            hipMalloc((void**)&n->concat_tmp, std::max<size_t>(max_concat, 16) * 4) == hipSuccess &&
            hipMalloc((void**)&n->part, max_part * 4) == hipSuccess &&
            hipMalloc((void**)&n->loss_part, piv::loss_part_floats(out_elems) * 4) == hipSuccess &&
            hipMalloc((void**)&n->loss, 16) == hipSuccess &&
            hipMalloc(&n->prep_scratch, piv::prepare_scratch_bytes(in_height * piv::kCnnFactor,
                                                                  in_width * piv::kCnnFactor, piv::kCnnFactor)) ==
                hipSuccess;
  for (int b = 1; b < kBuffers && ok; ++b) {
    const size_t elems = (size_t)max_batch * n->buf_h[b] * n->buf_w[b] * n->buf_c[b];
    if (b < kBuffers - 1) ok = hipMalloc((void**)&n->act[b], elems * 2) == hipSuccess;
    ok = ok && hipMalloc((void**)&n->gact[b], elems * 4) == hipSuccess;
  }
  if (!ok) {
    release(n);
    return capi_fail(PIV_E_NOMEM, "cnn: device allocation failed");
  }
  if (hipMemset(n->params, 0, off * 4) != hipSuccess || hipMemset(n->adam_m, 0, off * 4) != hipSuccess ||
      hipMemset(n->adam_v, 0, off * 4) != hipSuccess || hipMemset(n->grads, 0, off * 4) != hipSuccess ||
      hipMemset(n->wbf, 0, off * 2) != hipSuccess || hipMemset(n->wtbf, 0, off * 2) != hipSuccess) {
    release(n);
    return capi_fail(PIV_E_HIP, "cnn: hipMemset failed");
  }
  *out = n;
  return PIV_OK;
}

void piv_cnn_destroy(piv_cnn* net) { release(net); }

int piv_cnn_layout(const piv_cnn* n, int32_t* n_layers, int32_t* desc, int64_t* offsets, int64_t* n_params) {
  if (!n) return capi_fail(PIV_E_INVALID, "cnn: null network");
  if (n_layers) *n_layers = kLayers;
  if (n_params) *n_params = (int64_t)n->n_params;
  for (int l = 0; l < kLayers; ++l) {
    const Layer& L = n->L[l];
    if (desc) {
      int32_t* d = desc + 8 * l;
      d[0] = L.spec.s0;
      d[1] = L.spec.s1;
      d[2] = L.spec.up0;
      d[3] = L.c0;
      d[4] = L.c1;
      d[5] = L.spec.co;
      d[6] = L.spec.stride;
      d[7] = L.spec.relu;
    }
    if (offsets) {
      offsets[2 * l] = (int64_t)L.w_off;
      offsets[2 * l + 1] = (int64_t)L.b_off;
    }
  }
  return PIV_OK;
}

int piv_cnn_set_params(piv_cnn* n, const float* params) {
  if (!n || !params) return capi_fail(PIV_E_INVALID, "cnn: null argument");
  CNN_TRY(hipSetDevice(n->device));
  CNN_TRY(hipMemcpy(n->params, params, n->n_params * 4, hipMemcpyHostToDevice));
  CNN_TRY(hipMemset(n->adam_m, 0, n->n_params * 4));
  CNN_TRY(hipMemset(n->adam_v, 0, n->n_params * 4));
  n->step = 0;
  int r = convert_weights(n, nullptr);
  if (r) return r;
  CNN_TRY(hipDeviceSynchronize());
  return PIV_OK;
}

int piv_cnn_get_params(const piv_cnn* n, float* params) {
  if (!n || !params) return capi_fail(PIV_E_INVALID, "cnn: null argument");
  CNN_TRY(hipSetDevice(n->device));
  CNN_TRY(hipDeviceSynchronize());
  CNN_TRY(hipMemcpy(params, n->params, n->n_params * 4, hipMemcpyDeviceToHost));
  return PIV_OK;
}

int piv_cnn_get_grads(const piv_cnn* n, float* grads) {
  if (!n || !grads) return capi_fail(PIV_E_INVALID, "cnn: null argument");
  CNN_TRY(hipSetDevice(n->device));
  CNN_TRY(hipDeviceSynchronize());
  CNN_TRY(hipMemcpy(grads, n->grads, n->n_params * 4, hipMemcpyDeviceToHost));
  return PIV_OK;
}

int piv_cnn_get_buffer(const piv_cnn* n, int buffer, int gradient, int batch, float* out) {
  if (!n || !out || buffer < 1 || buffer >= kBuffers || batch < 1 || batch > n->max_batch)
    return capi_fail(PIV_E_INVALID, "cnn: bad buffer request");
  CNN_TRY(hipSetDevice(n->device));
  CNN_TRY(hipDeviceSynchronize());
  const size_t elems = (size_t)batch * n->buf_h[buffer] * n->buf_w[buffer] * n->buf_c[buffer];
  if (gradient) {
    CNN_TRY(hipMemcpy(out, n->gact[buffer], elems * 4, hipMemcpyDeviceToHost));
  } else if (buffer == kBuffers - 1) {
    CNN_TRY(hipMemcpy(out, n->out, elems * 4, hipMemcpyDeviceToHost));
  } else {
    std::vector<uint16_t> tmp(elems);
    CNN_TRY(hipMemcpy(tmp.data(), n->act[buffer], elems * 2, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < elems; ++i) {
      const uint32_t u = (uint32_t)tmp[i] << 16;
      memcpy(&out[i], &u, 4);
    }
  }
  return PIV_OK;
}

int piv_cnn_forward(piv_cnn* n, const uint16_t* input, int batch, float* out, void* stream) {
  int r = check_batch(n, batch);
  if (r) return r;
  if (!input) return capi_fail(PIV_E_INVALID, "cnn: null input");
  hipStream_t s = (hipStream_t)stream;
  r = forward(n, input, batch, s);
  if (r) return r;
  if (out)
    CNN_TRY(hipMemcpyAsync(out, n->out,
                           (size_t)batch * 2 * n->buf_h[kBuffers - 1] * n->buf_w[kBuffers - 1] * sizeof(float),
                           hipMemcpyDeviceToDevice, s));
  return PIV_OK;
}

int piv_cnn_gradients(piv_cnn* n, const uint16_t* input, const float* target, const float* weight, int batch,
                      float* loss, void* stream) {
  int r = check_batch(n, batch);
  if (r) return r;
  if (!input || !target) return capi_fail(PIV_E_INVALID, "cnn: null input or target");
  hipStream_t s = (hipStream_t)stream;
  r = forward(n, input, batch, s);
  if (r) return r;
  const long hw = (long)n->buf_h[kBuffers - 1] * n->buf_w[kBuffers - 1];
  const long count = (long)batch * 2 * hw;
  CNN_TRY(piv::launch_loss(n->out, target, weight, count, hw, n->gact[kBuffers - 1], n->loss_part,
                           loss ? loss : n->loss, s));
  return backward(n, input, batch, s);
}

int piv_cnn_adam_step(piv_cnn* n, float lr, float beta1, float beta2, float eps, void* stream) {
  if (!n) return capi_fail(PIV_E_INVALID, "cnn: null network");
  if (!(lr >= 0.f) || !(beta1 >= 0.f && beta1 < 1.f) || !(beta2 >= 0.f && beta2 < 1.f) || !(eps > 0.f))
    return capi_fail(PIV_E_INVALID, "cnn: bad Adam hyper-parameters");
  hipStream_t s = (hipStream_t)stream;
  n->step++;
  CNN_TRY(piv::launch_adam(n->params, n->grads, n->adam_m, n->adam_v, (long)n->n_params, lr, beta1, beta2, eps,
                           n->step, s));
  return convert_weights(n, s);
}

int piv_cnn_prepare_input(piv_cnn* n, const void* frame_a, const void* frame_b, int in_dtype, int height, int width,
                          uint16_t* out, void* stream) {
  if (!n || !frame_a || !frame_b || !out) return capi_fail(PIV_E_INVALID, "cnn: null argument");
  if (in_dtype < PIV_U8 || in_dtype > PIV_F32) return capi_fail(PIV_E_INVALID, "cnn: in_dtype");
  if (height != n->in_h * piv::kCnnFactor || width != n->in_w * piv::kCnnFactor)
    return capi_fail(PIV_E_INVALID, "cnn: frame %d x %d does not match the network input %d x %d (x%d)", height,
                     width, n->in_h, n->in_w, piv::kCnnFactor);
  CNN_TRY(piv::launch_prepare_input(frame_a, frame_b, in_dtype, height, width, piv::kCnnFactor, n->prep_scratch, out,
                                    (hipStream_t)stream));
  return PIV_OK;
}

}  // extern "C"
