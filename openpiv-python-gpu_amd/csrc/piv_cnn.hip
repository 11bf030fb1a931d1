// Learned coarse-displacement predictor: a small encoder-decoder CNN with
// hand-written training and inference kernels.  No counterpart in the
// reference, whose only predictor is the previous pass's field
// (gpu_process.py:620-625).
//
// Tensors are NHWC.  Activations and weights are BF16, every product
// accumulates in FP32 on v_mfma_f32_16x16x16_bf16; activation gradients,
// master weights and the Adam moments are FP32.
//
// Every 3x3 convolution (zero padding 1, stride 1 or 2) is an implicit GEMM:
//   forward      D[co][pixel]   = sum_{tap,ci} W[co][tap][ci] * X[in(pixel,tap)][ci]
//   input grad   D[ci][pixel]   = sum_{tap,co} W[co][tap][ci] * dY[out(pixel,tap)][co]
//   weight grad  D[co][tap,ci]  = sum_{pixel}  dY[pixel][co] * X[in(pixel,tap)][ci]
// A wave owns a 32 x 64 block of D (2 x 4 MFMA tiles).  MFMA operand layout
// (16x16x16): lane l holds A[l%16][4(l/16)+e] and B[4(l/16)+e][l%16], e = 0..3,
// and receives D[4(l/16)+e][l%16].  Decoder inputs are read through the
// nearest-neighbour x2 upsampling of one tensor concatenated with a skip
// tensor, so neither the upsampled nor the concatenated tensor exists.
#include <hip/hip_runtime.h>
#include <limits.h>
#include <math.h>
#include <stdint.h>

#include <algorithm>
#include <vector>

#include "../../include/openpiv_hip.h"
#include "piv_cnn.h"
#include "piv_kernels.h"

namespace piv {
int capi_fail(int code, const char* fmt, ...);
}

namespace piv {
namespace cnn {

typedef short bf16x4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float x) {
  uint32_t u = __builtin_bit_cast(uint32_t, x);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);  // NaN stays NaN
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __builtin_bit_cast(float, ((uint32_t)b) << 16); }

__device__ __forceinline__ bf16x4_t load_bf4(const uint16_t* p) {
  const uint2 r = *reinterpret_cast<const uint2*>(p);
  return __builtin_bit_cast(bf16x4_t, r);
}
__device__ __forceinline__ bf16x4_t f4_to_bf4(float4 f) {
  bf16x4_t r;
  r[0] = (short)f2bf(f.x);
  r[1] = (short)f2bf(f.y);
  r[2] = (short)f2bf(f.z);
  r[3] = (short)f2bf(f.w);
  return r;
}
__device__ __forceinline__ float4_t mfma(bf16x4_t a, bf16x4_t b, float4_t c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// Position (n, y, x) of flat pixel index p in a batch x h x w grid.
struct Pix {
  int n, y, x;
  bool ok;
};
__device__ __forceinline__ Pix pix_of(long p, long total, int h, int w) {
  Pix r;
  r.ok = p < total;
  const long q = r.ok ? p : 0;
  r.x = (int)(q % w);
  const long t = q / w;
  r.y = (int)(t % h);
  r.n = (int)(t / h);
  return r;
}

// Element offset of input channel 0 at conv-input position (n, yi, xi) in
// source 0 (possibly read through x2 upsampling) and source 1; -1 if the
// position is in the zero padding.
struct InView {
  const uint16_t* s0;
  const uint16_t* s1;
  int c0, c1, up0, hi, wi;
};
__device__ __forceinline__ bool in_offsets(const InView& v, int n, int yi, int xi, long* o0, long* o1) {
  if (yi < 0 || yi >= v.hi || xi < 0 || xi >= v.wi) return false;
  if (v.up0) {
    const int h0 = v.hi >> 1, w0 = v.wi >> 1;
    *o0 = (((long)n * h0 + (yi >> 1)) * w0 + (xi >> 1)) * v.c0;
  } else {
    *o0 = (((long)n * v.hi + yi) * v.wi + xi) * v.c0;
  }
  *o1 = (((long)n * v.hi + yi) * v.wi + xi) * v.c1;
  return true;
}
// 32-bit forms for the forward kernels (the launcher checks that every pixel
// index and element offset fits): 64-bit index arithmetic and division
// emulation were most of a small layer's instructions.
__device__ __forceinline__ Pix pix_of32(int p, int total, int h, int w) {
  Pix r;
  r.ok = p < total;
  const int q = r.ok ? p : 0;
  const int t = q / w;
  r.x = q - t * w;
  r.n = t / h;
  r.y = t - r.n * h;
  return r;
}
__device__ __forceinline__ bool in_offsets32(const InView& v, int n, int yi, int xi, int* o0, int* o1) {
  if (yi < 0 || yi >= v.hi || xi < 0 || xi >= v.wi) return false;
  if (v.up0) {
    const int h0 = v.hi >> 1, w0 = v.wi >> 1;
    *o0 = ((n * h0 + (yi >> 1)) * w0 + (xi >> 1)) * v.c0;
  } else {
    *o0 = ((n * v.hi + yi) * v.wi + xi) * v.c0;
  }
  *o1 = ((n * v.hi + yi) * v.wi + xi) * v.c1;
  return true;
}
__device__ __forceinline__ float in_value(const InView& v, bool ok, long o0, long o1, int ci) {
  if (!ok) return 0.f;
  return ci < v.c0 ? bf2f(v.s0[o0 + ci]) : bf2f(v.s1[o1 + ci - v.c0]);
}

// ---------------------------------------------------------------- forward
// One block computes one (16 TM) x 64 tile of D[co][pixel]; its four waves
// split the K = 9 * ci reduction (16-channel chunks of one tap, round robin)
// and are summed in LDS, so even the 32 x 32 layers of a single image spread
// over enough blocks.  ci0 and ci1 are multiples of 16 (4-channel vector loads).
// grid: x = ceil(pixels / 64), y = ceil(co / (16 TM)).
template <int TM, int NW = 4>
__global__ __launch_bounds__(64 * NW) void conv_fwd_kernel(ConvGeom g, InView in, const uint16_t* __restrict__ w,
                                                           const float* __restrict__ bias, void* out, int out_mode,
                                                           int relu) {
  __shared__ float4_t red[NW][TM][4][64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int r = lane & 15, gq = lane >> 4;
  const int total = g.batch * g.ho * g.wo;
  const int pbase = blockIdx.x * 64;
  const int cobase = blockIdx.y * (16 * TM);
  const int ci = in.c0 + in.c1, nc = ci / 16;
  Pix px[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) px[t] = pix_of32(pbase + t * 16 + r, total, g.ho, g.wo);
  int coA[TM];
  bool okA[TM];
#pragma unroll
  for (int m = 0; m < TM; ++m) {
    coA[m] = cobase + m * 16 + r;
    okA[m] = coA[m] < g.co;
  }
  float4_t acc[TM][4];
#pragma unroll
  for (int m = 0; m < TM; ++m)
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[m][t] = float4_t{0.f, 0.f, 0.f, 0.f};
  const bf16x4_t zero4 = {0, 0, 0, 0};
  // operands of K chunk kc: TM weight fragments and 4 input fragments
  auto load_chunk = [&](int kc, bf16x4_t (&a)[TM], bf16x4_t (&b)[4]) {
    const int tap = kc / nc, c = (kc - tap * nc) * 16;
    const int ky = tap / 3, kx = tap - 3 * ky;
    const int cc = c + 4 * gq;
#pragma unroll
    for (int m = 0; m < TM; ++m) a[m] = okA[m] ? load_bf4(w + ((coA[m] * 9 + tap) * ci + cc)) : zero4;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      int o0 = 0, o1 = 0;
      const bool ok = px[t].ok && in_offsets32(in, px[t].n, px[t].y * g.stride + ky - 1,
                                               px[t].x * g.stride + kx - 1, &o0, &o1);
      b[t] = !ok ? zero4 : (c < in.c0 ? load_bf4(in.s0 + (o0 + cc)) : load_bf4(in.s1 + (o1 + cc - in.c0)));
    }
  };
  // chunks kc = wv, wv + NW, ... of this wave, in groups of G whose loads are
  // all issued before the group's MFMAs (G chunks of latency in flight at once)
  constexpr int G = 4;
  const int K = 9 * nc;
  for (int kc0 = wv; kc0 < K; kc0 += NW * G) {
    bf16x4_t a[G][TM], b[G][4];
#pragma unroll
    for (int q = 0; q < G; ++q) {
      const int kc = kc0 + NW * q;
      if (kc < K) {
        load_chunk(kc, a[q], b[q]);
      } else {
#pragma unroll
        for (int m = 0; m < TM; ++m) a[q][m] = zero4;
#pragma unroll
        for (int t = 0; t < 4; ++t) b[q][t] = zero4;
      }
    }
#pragma unroll
    for (int q = 0; q < G; ++q)
#pragma unroll
      for (int m = 0; m < TM; ++m)
#pragma unroll
        for (int t = 0; t < 4; ++t) acc[m][t] = mfma(a[q][m], b[q][t], acc[m][t]);
  }
  // every wave's partial sums to LDS; wave w then sums and stores the
  // (m, t) tiles mt = 4m + t with mt mod NW = w, in wave order (the order of
  // a single accumulator walking the waves)
#pragma unroll
  for (int m = 0; m < TM; ++m)
#pragma unroll
    for (int t = 0; t < 4; ++t) red[wv][m][t][lane] = acc[m][t];
  __syncthreads();
#pragma unroll
  for (int mt = 0; mt < TM * 4; ++mt) {
    const int m = mt / 4, t = mt % 4;  // compile-time after unrolling: no indexed register arrays
    if (mt % NW != wv || !px[t].ok) continue;
    float4_t sum = red[0][m][t][lane];
#pragma unroll
    for (int q = 1; q < NW; ++q) sum += red[q][m][t][lane];
    // rows co = cobase + 16m + 4gq + e of pixel column r
    const int p = pbase + t * 16 + r;
    const int co0 = cobase + m * 16 + 4 * gq;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int co = co0 + e;
      if (co >= g.co) continue;
      float val = sum[e] + bias[co];
      if (relu) val = fmaxf(val, 0.f);
      if (out_mode == kOutBf16) {
        static_cast<uint16_t*>(out)[p * g.co + co] = f2bf(val);
      } else if (out_mode == kOutF32) {
        static_cast<float*>(out)[p * g.co + co] = val;
      } else {  // per-image planar [n][co][y][x]
        const int hw = g.ho * g.wo;
        static_cast<float*>(out)[(px[t].n * g.co + co) * hw + (px[t].y * g.wo + px[t].x)] = val;
      }
    }
  }
}

// Layers whose input channels are not multiples of 16 (the 2-channel input
// layer): each wave owns a 32 x (16 TP) tile and walks the flat (tap, ci)
// index element by element.  The layer is a few MFLOP over a full-resolution
// grid: one 16-pixel tile per wave keeps four times more waves in flight
// than four tiles would, which is what its load latency needs.
// grid: x = ceil(pixels / (64 TP)), y = ceil(co / 32).
constexpr int kSmallTP = 1;
template <int CI>  // CI > 0: channel count known at compile time
__global__ __launch_bounds__(256) void conv_fwd_small_kernel(ConvGeom g, InView in, const uint16_t* __restrict__ w,
                                                             const float* __restrict__ bias, void* out, int out_mode,
                                                             int relu) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int r = lane & 15, gq = lane >> 4;
  const int total = g.batch * g.ho * g.wo;
  const int pbase = (blockIdx.x * 4 + wv) * (16 * kSmallTP);
  if (pbase >= total) return;
  const int cobase = blockIdx.y * 32;
  const int ci = CI > 0 ? CI : in.c0 + in.c1, K = 9 * ci;
  Pix px[kSmallTP];
#pragma unroll
  for (int t = 0; t < kSmallTP; ++t) px[t] = pix_of32(pbase + t * 16 + r, total, g.ho, g.wo);
  int coA[2];
  bool okA[2];
#pragma unroll
  for (int m = 0; m < 2; ++m) {
    coA[m] = cobase + m * 16 + r;
    okA[m] = coA[m] < g.co;
  }
  float4_t acc[2][kSmallTP];
#pragma unroll
  for (int m = 0; m < 2; ++m)
#pragma unroll
    for (int t = 0; t < kSmallTP; ++t) acc[m][t] = float4_t{0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += 16) {
    bf16x4_t a[2], b[kSmallTP];
#pragma unroll
    for (int m = 0; m < 2; ++m)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int k = k0 + 4 * gq + e;
        a[m][e] = (okA[m] && k < K) ? (short)w[coA[m] * K + k] : (short)0;
      }
#pragma unroll
    for (int t = 0; t < kSmallTP; ++t)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int k = k0 + 4 * gq + e;
        float val = 0.f;
        if (k < K && px[t].ok) {
          const int tap = k / ci, cch = k - tap * ci;
          int o0 = 0, o1 = 0;
          const bool ok = in_offsets32(in, px[t].n, px[t].y * g.stride + tap / 3 - 1,
                                       px[t].x * g.stride + tap % 3 - 1, &o0, &o1);
          val = in_value(in, ok, o0, o1, cch);
        }
        b[t][e] = (short)f2bf(val);
      }
#pragma unroll
    for (int m = 0; m < 2; ++m)
#pragma unroll
      for (int t = 0; t < kSmallTP; ++t) acc[m][t] = mfma(a[m], b[t], acc[m][t]);
  }
#pragma unroll
  for (int t = 0; t < kSmallTP; ++t) {
    if (!px[t].ok) continue;
    const int p = pbase + t * 16 + r;
#pragma unroll
    for (int m = 0; m < 2; ++m) {
      const int co0 = cobase + m * 16 + 4 * gq;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int co = co0 + e;
        if (co >= g.co) continue;
        float val = acc[m][t][e] + bias[co];
        if (relu) val = fmaxf(val, 0.f);
        if (out_mode == kOutBf16) {
          static_cast<uint16_t*>(out)[p * g.co + co] = f2bf(val);
        } else if (out_mode == kOutF32) {
          static_cast<float*>(out)[p * g.co + co] = val;
        } else {
          const int hw = g.ho * g.wo;
          static_cast<float*>(out)[(px[t].n * g.co + co) * hw + (px[t].y * g.wo + px[t].x)] = val;
        }
      }
    }
  }
}

// ------------------------------------------------------------ input grad
// D[ci][in-pixel]; grid x = ceil(in-pixels / 256), y = ceil(ci / 32).
// dy: FP32 [batch][ho][wo][co] (already multiplied by the ReLU mask).
// wt: BF16 [ci][tap][co].  dx: FP32 [batch][hi][wi][ci]; accumulate adds.
template <bool VEC>
__global__ __launch_bounds__(256) void conv_bwd_data_kernel(ConvGeom g, int ci, const float* __restrict__ dy,
                                                            const uint16_t* __restrict__ wt, float* dx,
                                                            int accumulate) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int r = lane & 15, gq = lane >> 4;
  const long total = (long)g.batch * g.hi * g.wi;
  const long pbase = ((long)blockIdx.x * 4 + wv) * 64;
  if (pbase >= total) return;
  const int cibase = blockIdx.y * 32;
  const int co = g.co, K = 9 * co;
  Pix px[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) px[t] = pix_of(pbase + t * 16 + r, total, g.hi, g.wi);
  int ciA[2];
  bool okA[2];
#pragma unroll
  for (int m = 0; m < 2; ++m) {
    ciA[m] = cibase + m * 16 + r;
    okA[m] = ciA[m] < ci;
  }
  float4_t acc[2][4];
#pragma unroll
  for (int m = 0; m < 2; ++m)
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[m][t] = float4_t{0.f, 0.f, 0.f, 0.f};
  const bf16x4_t zero4 = {0, 0, 0, 0};
  // output position fed by input (yi, xi) through tap (ky, kx): s*yo + ky - 1 = yi
  auto out_off = [&](const Pix& q, int tap, long* off) -> bool {
    if (!q.ok) return false;
    const int ny = q.y + 1 - tap / 3, nx = q.x + 1 - tap % 3;
    if (ny < 0 || nx < 0) return false;
    if (g.stride == 2 && ((ny | nx) & 1)) return false;
    const int yo = ny / g.stride, xo = nx / g.stride;
    if (yo >= g.ho || xo >= g.wo) return false;
    *off = (((long)q.n * g.ho + yo) * g.wo + xo) * co;
    return true;
  };
  if constexpr (VEC) {
    for (int tap = 0; tap < 9; ++tap) {
      long off[4];
      bool ok[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) ok[t] = out_off(px[t], tap, &off[t]);
      for (int c = 0; c < co; c += 16) {
        const int cc = c + 4 * gq;
        bf16x4_t a[2], b[4];
#pragma unroll
        for (int m = 0; m < 2; ++m) a[m] = okA[m] ? load_bf4(wt + ((long)ciA[m] * 9 + tap) * co + cc) : zero4;
#pragma unroll
        for (int t = 0; t < 4; ++t)
          b[t] = ok[t] ? f4_to_bf4(*reinterpret_cast<const float4*>(dy + off[t] + cc)) : zero4;
#pragma unroll
        for (int m = 0; m < 2; ++m)
#pragma unroll
          for (int t = 0; t < 4; ++t) acc[m][t] = mfma(a[m], b[t], acc[m][t]);
      }
    }
  } else {
    for (int k0 = 0; k0 < K; k0 += 16) {
      bf16x4_t a[2], b[4];
#pragma unroll
      for (int m = 0; m < 2; ++m)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int k = k0 + 4 * gq + e;
          a[m][e] = (okA[m] && k < K) ? (short)wt[(long)ciA[m] * K + k] : (short)0;
        }
#pragma unroll
      for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int k = k0 + 4 * gq + e;
          float val = 0.f;
          long off;
          if (k < K && out_off(px[t], k / co, &off)) val = dy[off + k % co];
          b[t][e] = (short)f2bf(val);
        }
#pragma unroll
      for (int m = 0; m < 2; ++m)
#pragma unroll
        for (int t = 0; t < 4; ++t) acc[m][t] = mfma(a[m], b[t], acc[m][t]);
    }
  }
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    if (!px[t].ok) continue;
    const long p = pbase + t * 16 + r;
#pragma unroll
    for (int m = 0; m < 2; ++m) {
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int c = cibase + m * 16 + 4 * gq + e;
        if (c >= ci) continue;
        float* d = dx + p * ci + c;
        *d = accumulate ? *d + acc[m][t][e] : acc[m][t][e];
      }
    }
  }
}

// ----------------------------------------------------------- weight grad
// Partial sums over one chunk of output pixels:
//   part[z][co][n] = sum_{pixel in chunk z} dy[pixel][co] * X[in(pixel, tap)][ci]
// for n = tap * ci_total + ci < 9 ci_total, and for the extra column
// n = 9 ci_total the bias gradient sum dy[pixel][co] (an all-ones operand).
// grid x = ceil((9 ci + 1) / 64), y = ceil(co / 32), z = chunks; the four
// waves of a block split the chunk and are summed in LDS.
__global__ __launch_bounds__(256) void conv_bwd_weight_kernel(ConvGeom g, InView in, const float* __restrict__ dy,
                                                              float* part, int chunk_px) {
  __shared__ float red[3][2][4][64][4];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int r = lane & 15, gq = lane >> 4;
  const int ci = in.c0 + in.c1, NW = 9 * ci, N = NW + 1;
  const long total = (long)g.batch * g.ho * g.wo;
  const long p0 = (long)blockIdx.z * chunk_px;
  const long p1 = min(total, p0 + chunk_px);
  const int nbase = blockIdx.x * 64, cobase = blockIdx.y * 32;
  int coA[2];
  bool okA[2];
#pragma unroll
  for (int m = 0; m < 2; ++m) {
    coA[m] = cobase + m * 16 + r;
    okA[m] = coA[m] < g.co;
  }
  int tapB[4], ciB[4], kindB[4];  // kind: 0 outside, 1 weight column, 2 bias column
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int n = nbase + t * 16 + r;
    kindB[t] = n < NW ? 1 : (n == NW ? 2 : 0);
    tapB[t] = n < NW ? n / ci : 0;
    ciB[t] = n < NW ? n % ci : 0;
  }
  float4_t acc[2][4];
#pragma unroll
  for (int m = 0; m < 2; ++m)
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[m][t] = float4_t{0.f, 0.f, 0.f, 0.f};
  for (long k0 = p0 + wv * 16; k0 < p1; k0 += 64) {
    bf16x4_t a[2], b[4];
    Pix q[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) q[e] = pix_of(k0 + 4 * gq + e, p1, g.ho, g.wo);
#pragma unroll
    for (int m = 0; m < 2; ++m)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const long p = k0 + 4 * gq + e;
        a[m][e] = (short)f2bf((okA[m] && q[e].ok) ? dy[p * g.co + coA[m]] : 0.f);
      }
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        float val = 0.f;
        if (q[e].ok && kindB[t] == 2) {
          val = 1.f;
        } else if (q[e].ok && kindB[t] == 1) {
          long o0, o1;
          const bool ok = in_offsets(in, q[e].n, q[e].y * g.stride + tapB[t] / 3 - 1,
                                     q[e].x * g.stride + tapB[t] % 3 - 1, &o0, &o1);
          val = in_value(in, ok, o0, o1, ciB[t]);
        }
        b[t][e] = (short)f2bf(val);
      }
#pragma unroll
    for (int m = 0; m < 2; ++m)
#pragma unroll
      for (int t = 0; t < 4; ++t) acc[m][t] = mfma(a[m], b[t], acc[m][t]);
  }
  if (wv > 0) {
#pragma unroll
    for (int m = 0; m < 2; ++m)
#pragma unroll
      for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int e = 0; e < 4; ++e) red[wv - 1][m][t][lane][e] = acc[m][t][e];
  }
  __syncthreads();
  if (wv != 0) return;
  float* out = part + (long)blockIdx.z * g.co * N;
#pragma unroll
  for (int m = 0; m < 2; ++m)
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int n = nbase + t * 16 + r;
      if (n >= N) continue;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int c = cobase + m * 16 + 4 * gq + e;
        if (c >= g.co) continue;
        const float s = acc[m][t][e] + red[0][m][t][lane][e] + red[1][m][t][lane][e] + red[2][m][t][lane][e];
        out[(long)c * N + n] = s;
      }
    }
}

// dw[c][n] / db[c] = sum_z part[z][c][n] in z order (deterministic);
// column n = nw of the partials is the bias.
__global__ __launch_bounds__(256) void reduce_wgrad_kernel(const float* __restrict__ part, int nz, int co, int nw,
                                                           float* dw, float* db) {
  const long count = (long)co * (nw + 1);
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  float s = 0.f;
  for (int z = 0; z < nz; ++z) s += part[(long)z * count + i];
  const int c = (int)(i / (nw + 1)), n = (int)(i % (nw + 1));
  if (n < nw)
    dw[(long)c * nw + n] = s;
  else
    db[c] = s;
}

// g *= (activation > 0): ReLU backward on a BF16 activation.
__global__ __launch_bounds__(256) void relu_mask_kernel(float* g, const uint16_t* __restrict__ act, long count) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const uint16_t b = act[i];
  if ((b & 0x8000u) || b == 0) g[i] = 0.f;
}

// Gradient of the concatenated, upsampled decoder input [batch][hi][wi][c0+c1]
// split into its sources: d0 [batch][hi/2][wi/2][c0] = sum of the 2x2 block
// (nearest-neighbour upsampling), d1 [batch][hi][wi][c1] = the skip part.
__global__ __launch_bounds__(256) void split_grad_kernel(const float* __restrict__ t, int batch, int hi, int wi,
                                                         int c0, int c1, float* d0, float* d1) {
  const long n0 = (long)batch * (hi / 2) * (wi / 2) * c0;
  const long n1 = (long)batch * hi * wi * c1;
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int ct = c0 + c1;
  if (i < n0) {
    const int c = (int)(i % c0);
    long q = i / c0;
    const int x = (int)(q % (wi / 2));
    q /= (wi / 2);
    const int y = (int)(q % (hi / 2));
    const int n = (int)(q / (hi / 2));
    float s = 0.f;
#pragma unroll
    for (int dy = 0; dy < 2; ++dy)
#pragma unroll
      for (int dx = 0; dx < 2; ++dx) s += t[(((long)n * hi + 2 * y + dy) * wi + 2 * x + dx) * ct + c];
    d0[i] = s;
  } else if (i < n0 + n1) {
    const long j = i - n0;
    const int c = (int)(j % c1);
    const long p = j / c1;
    d1[j] = t[p * ct + c0 + c];
  }
}

// Weighted mean-square loss over a per-image planar output [batch][2][h][w]:
//   L = sum wt (o - y)^2 / count, dL/do = 2 wt (o - y) / count,
// wt per vector [batch][h][w] (both components).  The gradient is written in
// NHWC [batch][h][w][2] for the head's backward; part[block] = the block's
// sum of wt (o - y)^2.
__global__ __launch_bounds__(256) void loss_kernel(const float* __restrict__ out, const float* __restrict__ target,
                                                   const float* __restrict__ weight, long count, long hw, float* grad,
                                                   float* part) {
  __shared__ float red[256];
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  float s = 0.f;
  if (i < count) {
    const long n = i / (2 * hw), c = (i / hw) & 1, q = i % hw;
    const float wgt = weight ? weight[n * hw + q] : 1.f;
    const float d = out[i] - target[i];
    grad[(n * hw + q) * 2 + c] = 2.f * wgt * d / (float)count;
    s = wgt * d * d;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void loss_final_kernel(const float* __restrict__ part, int nparts, long count, float* loss) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double s = 0.0;
  for (int i = 0; i < nparts; ++i) s += part[i];
  *loss = (float)(s / (double)count);
}

__global__ __launch_bounds__(256) void adam_kernel(float* p, const float* __restrict__ g, float* m, float* v,
                                                   long count, float lr, float b1, float b2, float eps, float c1,
                                                   float c2) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// Master FP32 weights [co][tap][ci] -> BF16 copies in that layout and
// transposed [ci][tap][co] (the input-gradient operand).
__global__ __launch_bounds__(256) void weight_convert_kernel(const float* __restrict__ p, int co, int ci,
                                                             uint16_t* w, uint16_t* wt) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long count = (long)co * 9 * ci;
  if (i >= count) return;
  const int c = (int)(i % ci);
  const int tap = (int)((i / ci) % 9);
  const int o = (int)(i / (9L * ci));
  const uint16_t b = f2bf(p[i]);
  w[i] = b;
  wt[((long)c * 9 + tap) * co + o] = b;
}

// Box means of each 8 x 8 block of a frame -> tmp[frame][y][x] (FP32), with
// per-block partial sums for the frame statistics.
template <typename T>
__global__ __launch_bounds__(256) void box_mean_kernel(const T* __restrict__ fa, const T* __restrict__ fb, int ht,
                                                       int wd, int f, float* tmp, double* part) {
  __shared__ double red[2][256];
  const int hh = ht / f, ww = wd / f;
  const int frame = blockIdx.y;
  const T* src = frame ? fb : fa;
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double s = 0.0, s2 = 0.0;
  if (i < (long)hh * ww) {
    const int y = (int)(i / ww), x = (int)(i % ww);
    float acc = 0.f;
    if constexpr (sizeof(T) == 2) {
      if (f == 8 && (wd & 7) == 0) {  // one 16-byte load per row
        for (int dy = 0; dy < 8; ++dy) {
          const uint4 q = *reinterpret_cast<const uint4*>(src + (long)(y * 8 + dy) * wd + x * 8);
          const uint32_t wds[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
          for (int k = 0; k < 4; ++k) acc += (float)(wds[k] & 0xffffu) + (float)(wds[k] >> 16);
        }
      } else {
        for (int dy = 0; dy < f; ++dy) {
          const T* row = src + (long)(y * f + dy) * wd + x * f;
          for (int dx = 0; dx < f; ++dx) acc += (float)row[dx];
        }
      }
    } else {
      for (int dy = 0; dy < f; ++dy) {
        const T* row = src + (long)(y * f + dy) * wd + x * f;
        for (int dx = 0; dx < f; ++dx) acc += (float)row[dx];
      }
    }
    acc /= (float)(f * f);
    tmp[(long)frame * hh * ww + i] = acc;
    s = acc;
    s2 = (double)acc * acc;
  }
  red[0][threadIdx.x] = s;
  red[1][threadIdx.x] = s2;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      red[0][threadIdx.x] += red[0][threadIdx.x + o];
      red[1][threadIdx.x] += red[1][threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[((long)frame * gridDim.x + blockIdx.x) * 2] = red[0][0];
    part[((long)frame * gridDim.x + blockIdx.x) * 2 + 1] = red[1][0];
  }
}

// Standardise each frame's box means (zero mean, unit variance) into the
// BF16 NHWC input [y][x][frame].
__global__ __launch_bounds__(256) void standardise_kernel(const float* __restrict__ tmp, const double* __restrict__ part,
                                                          int nparts, long npix, uint16_t* out) {
  __shared__ double red[4][256];
  __shared__ float stat[2][2];
  double a[4] = {0.0, 0.0, 0.0, 0.0};  // sum, sum of squares of frame 0 and frame 1
  for (int k = threadIdx.x; k < nparts; k += blockDim.x) {
    a[0] += part[(long)k * 2];
    a[1] += part[(long)k * 2 + 1];
    a[2] += part[((long)nparts + k) * 2];
    a[3] += part[((long)nparts + k) * 2 + 1];
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) red[q][threadIdx.x] = a[q];
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
#pragma unroll
      for (int q = 0; q < 4; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x < 2) {
    const int frame = threadIdx.x;
    const double mean = red[2 * frame][0] / (double)npix;
    const double var = fmax(red[2 * frame + 1][0] / (double)npix - mean * mean, 0.0);
    stat[frame][0] = (float)mean;
    stat[frame][1] = (float)(var > 0.0 ? 1.0 / sqrt(var) : 1.0);
  }
  __syncthreads();
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npix) return;
#pragma unroll
  for (int frame = 0; frame < 2; ++frame)
    out[i * 2 + frame] = f2bf((tmp[(long)frame * npix + i] - stat[frame][0]) * stat[frame][1]);
}

inline unsigned blocks_for(long n, int b = 256) { return (unsigned)((n + b - 1) / b); }

}  // namespace cnn

using namespace cnn;

hipError_t launch_conv_fwd(const ConvGeom& g, const uint16_t* s0, int c0, int up0, const uint16_t* s1, int c1,
                           const uint16_t* w, const float* bias, void* out, int out_mode, int relu, hipStream_t s) {
  const InView in{s0, s1, c0, c1, up0, g.hi, g.wi};
  const long pixels = (long)g.batch * g.ho * g.wo;
  // the forward kernels index in 32 bits
  if ((long)g.batch * g.hi * g.wi * (c0 + c1 + 1) >= INT_MAX || (pixels + 256) * (g.co + 32) >= INT_MAX ||
      (long)g.co * 9 * (c0 + c1) >= INT_MAX)
    return hipErrorInvalidValue;
  if (c0 % 16 == 0 && c1 % 16 == 0) {
    // 32-row tiles once there are enough 64-pixel columns to fill the chip
    const long cols = (pixels + 63) / 64;
    if (g.co > 16 && cols * ((g.co + 31) / 32) >= 512) {
      const dim3 grid((unsigned)cols, (unsigned)((g.co + 31) / 32));
      if (cols * ((g.co + 31) / 32) < 1024)
        hipLaunchKernelGGL((conv_fwd_kernel<2, 8>), grid, dim3(512), 0, s, g, in, w, bias, out, out_mode, relu);
      else
        hipLaunchKernelGGL(conv_fwd_kernel<2>, grid, dim3(256), 0, s, g, in, w, bias, out, out_mode, relu);
    } else {
      // fewer tiles than CUs x 2: eight waves split each tile's K
      const dim3 grid((unsigned)cols, (unsigned)((g.co + 15) / 16));
      if (cols * ((g.co + 15) / 16) < 512)
        hipLaunchKernelGGL((conv_fwd_kernel<1, 8>), grid, dim3(512), 0, s, g, in, w, bias, out, out_mode, relu);
      else
        hipLaunchKernelGGL(conv_fwd_kernel<1>, grid, dim3(256), 0, s, g, in, w, bias, out, out_mode, relu);
    }
  } else {
    const dim3 grid(blocks_for(pixels, 64 * kSmallTP), (unsigned)((g.co + 31) / 32));
    if (c0 + c1 == 2)
      hipLaunchKernelGGL(conv_fwd_small_kernel<2>, grid, dim3(256), 0, s, g, in, w, bias, out, out_mode, relu);
    else
      hipLaunchKernelGGL(conv_fwd_small_kernel<0>, grid, dim3(256), 0, s, g, in, w, bias, out, out_mode, relu);
  }
  return hipGetLastError();
}

hipError_t launch_conv_bwd_data(const ConvGeom& g, int ci, const float* dy, const uint16_t* wt, float* dx,
                                int accumulate, hipStream_t s) {
  const long pixels = (long)g.batch * g.hi * g.wi;
  const dim3 grid(blocks_for(pixels), (unsigned)((ci + 31) / 32));
  if (g.co % 16 == 0)
    hipLaunchKernelGGL(conv_bwd_data_kernel<true>, grid, dim3(256), 0, s, g, ci, dy, wt, dx, accumulate);
  else
    hipLaunchKernelGGL(conv_bwd_data_kernel<false>, grid, dim3(256), 0, s, g, ci, dy, wt, dx, accumulate);
  return hipGetLastError();
}

// Pixels per weight-gradient chunk: at least 1024, and at most 256 chunks
// (chunk >= pixels / 256), so a call has nz <= min(256, ceil(pixels / 1024)).
int weight_grad_chunk(const ConvGeom& g) {
  const long pixels = (long)g.batch * g.ho * g.wo;
  long chunk = std::max<long>(1024, (pixels + 255) / 256);
  chunk = (chunk + 63) / 64 * 64;
  return (int)chunk;
}

// Workspace for any batch up to g.batch: the chunk-count bound above is
// non-decreasing in the pixel count.
size_t weight_grad_part_floats(const ConvGeom& g, int ci) {
  const long pixels = (long)g.batch * g.ho * g.wo;
  const long nz = std::min<long>(256, (pixels + 1023) / 1024);
  return (size_t)nz * g.co * (9 * ci + 1);
}

hipError_t launch_conv_bwd_weight(const ConvGeom& g, const uint16_t* s0, int c0, int up0, const uint16_t* s1, int c1,
                                  const float* dy, float* part, float* dw, float* db, hipStream_t s) {
  const InView in{s0, s1, c0, c1, up0, g.hi, g.wi};
  const int ci = c0 + c1;
  const long pixels = (long)g.batch * g.ho * g.wo;
  const int chunk = weight_grad_chunk(g);
  const int nz = (int)((pixels + chunk - 1) / chunk);
  const dim3 grid((unsigned)((9 * ci + 1 + 63) / 64), (unsigned)((g.co + 31) / 32), (unsigned)nz);
  hipLaunchKernelGGL(conv_bwd_weight_kernel, grid, dim3(256), 0, s, g, in, dy, part, chunk);
  const long count = (long)g.co * (9 * ci + 1);
  hipLaunchKernelGGL(reduce_wgrad_kernel, dim3(blocks_for(count)), dim3(256), 0, s, part, nz, g.co, 9 * ci, dw, db);
  return hipGetLastError();
}

hipError_t launch_relu_mask(float* g, const uint16_t* act, long count, hipStream_t s) {
  hipLaunchKernelGGL(relu_mask_kernel, dim3(blocks_for(count)), dim3(256), 0, s, g, act, count);
  return hipGetLastError();
}

hipError_t launch_split_grad(const float* t, int batch, int hi, int wi, int c0, int c1, float* d0, float* d1,
                             hipStream_t s) {
  const long n = (long)batch * (hi / 2) * (wi / 2) * c0 + (long)batch * hi * wi * c1;
  hipLaunchKernelGGL(split_grad_kernel, dim3(blocks_for(n)), dim3(256), 0, s, t, batch, hi, wi, c0, c1, d0, d1);
  return hipGetLastError();
}

size_t loss_part_floats(long count) { return blocks_for(count); }

hipError_t launch_loss(const float* out, const float* target, const float* weight, long count, long hw, float* grad,
                       float* part, float* loss, hipStream_t s) {
  const unsigned nb = blocks_for(count);
  hipLaunchKernelGGL(loss_kernel, dim3(nb), dim3(256), 0, s, out, target, weight, count, hw, grad, part);
  if (loss) hipLaunchKernelGGL(loss_final_kernel, dim3(1), dim3(64), 0, s, part, (int)nb, count, loss);
  return hipGetLastError();
}

hipError_t launch_adam(float* p, const float* g, float* m, float* v, long count, float lr, float b1, float b2,
                       float eps, int step, hipStream_t s) {
  const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
  hipLaunchKernelGGL(adam_kernel, dim3(blocks_for(count)), dim3(256), 0, s, p, g, m, v, count, lr, b1, b2, eps, c1,
                     c2);
  return hipGetLastError();
}

hipError_t launch_weight_convert(const float* p, int co, int ci, uint16_t* w, uint16_t* wt, hipStream_t s) {
  const long count = (long)co * 9 * ci;
  hipLaunchKernelGGL(weight_convert_kernel, dim3(blocks_for(count)), dim3(256), 0, s, p, co, ci, w, wt);
  return hipGetLastError();
}

size_t prepare_scratch_bytes(int ht, int wd, int f) {
  const long npix = (long)(ht / f) * (wd / f);
  return sizeof(float) * 2 * npix + sizeof(double) * 2 * 2 * blocks_for(npix) + 256;
}

hipError_t launch_prepare_input(const void* fa, const void* fb, int dtype, int ht, int wd, int f, void* scratch,
                                uint16_t* out, hipStream_t s) {
  const long npix = (long)(ht / f) * (wd / f);
  const unsigned nb = blocks_for(npix);
  float* tmp = static_cast<float*>(scratch);
  double* part = reinterpret_cast<double*>(static_cast<char*>(scratch) + ((sizeof(float) * 2 * npix + 255) / 256) * 256);
  const dim3 grid(nb, 2);
  if (dtype == kU8)
    hipLaunchKernelGGL(box_mean_kernel<uint8_t>, grid, dim3(256), 0, s, (const uint8_t*)fa, (const uint8_t*)fb, ht, wd,
                       f, tmp, part);
  else if (dtype == kU16)
    hipLaunchKernelGGL(box_mean_kernel<uint16_t>, grid, dim3(256), 0, s, (const uint16_t*)fa, (const uint16_t*)fb, ht,
                       wd, f, tmp, part);
  else
    hipLaunchKernelGGL(box_mean_kernel<float>, grid, dim3(256), 0, s, (const float*)fa, (const float*)fb, ht, wd, f,
                       tmp, part);
  hipLaunchKernelGGL(standardise_kernel, dim3(nb), dim3(256), 0, s, tmp, part, (int)nb, npix, out);
  return hipGetLastError();
}

}  // namespace piv
