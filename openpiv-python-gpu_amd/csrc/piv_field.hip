// Field-level kernels of one PIV pass: frame deformation, grid-to-grid
// predictor interpolation, 3x3 median validation and spring replacement.
#include <math.h>
#include <stdio.h>
#include <algorithm>

#include "piv_buffer.h"
#include "piv_fft.h"
#include "piv_kernels.h"
#include "piv_record.h"
#include "piv_wave.h"

namespace piv {

namespace {

template <typename T>
__device__ __forceinline__ float ld(const void* base, size_t idx) {
  return (float)(reinterpret_cast<const T*>(base)[idx]);
}
template <typename R, typename T>
__device__ __forceinline__ R ldr(const void* base, size_t idx) {
  return (R)(reinterpret_cast<const T*>(base)[idx]);
}

// Precision-generic scalar helpers (float and double).
__device__ __forceinline__ float vfma(float a, float b, float c) { return fmaf(a, b, c); }
__device__ __forceinline__ double vfma(double a, double b, double c) { return fma(a, b, c); }
__device__ __forceinline__ float vfloor(float a) { return floorf(a); }
__device__ __forceinline__ double vfloor(double a) { return floor(a); }
__device__ __forceinline__ float vabs(float a) { return fabsf(a); }
__device__ __forceinline__ double vabs(double a) { return fabs(a); }
__device__ __forceinline__ float vmin(float a, float b) { return fminf(a, b); }
__device__ __forceinline__ double vmin(double a, double b) { return fmin(a, b); }
__device__ __forceinline__ float vmax(float a, float b) { return fmaxf(a, b); }
__device__ __forceinline__ double vmax(double a, double b) { return fmax(a, b); }
template <typename R> __device__ __forceinline__ R vnan();
template <> __device__ __forceinline__ float vnan<float>() { return __builtin_nanf(""); }
template <> __device__ __forceinline__ double vnan<double>() { return __builtin_nan(""); }

// Bilinear value of a regular-grid field at (yq, xq) with linear
// extrapolation outside the grid: the interval index is clamped to
// [0, n-2] while the normalised distance is not (cupyx/scipy
// RegularGridInterpolator(method="linear", fill_value=None) semantics used at
// gpu_process.py:737-740 and :1159-1162).
// Grid cell (clamped) of a coordinate along one axis.
template <typename R>
__device__ __forceinline__ int grid_cell(R q, R q0, R inv, int cells) {
  const int i = (int)vfloor((q - q0) * inv);
  return i < 0 ? 0 : (i > cells - 2 ? cells - 2 : i);
}

template <typename R>
__device__ __forceinline__ void grid_bilinear(const R* fu, const R* fv, int m, int n, R y0, R x0, R sp, R yq, R xq,
                                              R& ou, R& ov) {
  // one division, then products (exact for the power-of-two spacings of the
  // usual schedules; otherwise within a rounding of the quotient)
  const R inv = R(1) / sp;
  const int iy = grid_cell(yq, y0, inv, m);
  const int ix = grid_cell(xq, x0, inv, n);
  const R ty = (yq - (y0 + (R)iy * sp)) * inv;
  const R tx = (xq - (x0 + (R)ix * sp)) * inv;
  const R wy0 = R(1) - ty, wx0 = R(1) - tx;
  const size_t i00 = (size_t)iy * n + ix;
  const size_t i10 = i00 + n;
  ou = fu[i00] * wy0 * wx0 + fu[i00 + 1] * wy0 * tx + fu[i10] * ty * wx0 + fu[i10 + 1] * ty * tx;
  ov = fv[i00] * wy0 * wx0 + fv[i00 + 1] * wy0 * tx + fv[i10] * ty * wx0 + fv[i10 + 1] * ty * tx;
}

// grid_bilinear with the grid read from an LDS patch holding grid rows
// iy_lo.. and columns ix_lo.. (row stride PS).
template <typename R, int PS>
__device__ __forceinline__ void grid_bilinear_lds(const R* su, const R* sv, int iy_lo, int ix_lo, int m, int n, R y0,
                                                  R x0, R sp, R yq, R xq, R& ou, R& ov) {
  const R inv = R(1) / sp;
  const int iy = grid_cell(yq, y0, inv, m);
  const int ix = grid_cell(xq, x0, inv, n);
  const R ty = (yq - (y0 + (R)iy * sp)) * inv;
  const R tx = (xq - (x0 + (R)ix * sp)) * inv;
  const R wy0 = R(1) - ty, wx0 = R(1) - tx;
  const int i00 = (iy - iy_lo) * PS + (ix - ix_lo);
  const int i10 = i00 + PS;
  ou = su[i00] * wy0 * wx0 + su[i00 + 1] * wy0 * tx + su[i10] * ty * wx0 + su[i10 + 1] * ty * tx;
  ov = sv[i00] * wy0 * wx0 + sv[i00 + 1] * wy0 * tx + sv[i10] * ty * wx0 + sv[i10 + 1] * ty * tx;
}

// cuda_interpolate_frame (gpu_process.py:1528-1546): clamped base index,
// unclamped fraction, so the border extrapolates linearly.
template <typename R, typename T>
__device__ __forceinline__ R warp_px(const void* f, int ht, int wd, R X, R Y, R up, R vp, R dt) {
  const R xi = vfma(up, dt, X);
  const R yi = vfma(vp, dt, Y);
  int xl = (int)vfloor(xi);
  int yu = (int)vfloor(yi);
  yu = yu < 0 ? 0 : (yu > ht - 2 ? ht - 2 : yu);
  xl = xl < 0 ? 0 : (xl > wd - 2 ? wd - 2 : xl);
  const int yd = yu + 1, xr = xl + 1;
  const R ful = ldr<R, T>(f, (size_t)yu * wd + xl);
  const R fur = ldr<R, T>(f, (size_t)yu * wd + xr);
  const R fdl = ldr<R, T>(f, (size_t)yd * wd + xl);
  const R fdr = ldr<R, T>(f, (size_t)yd * wd + xr);
  const R a0 = yi - (R)yu, a1 = (R)yd - yi;
  const R b0 = (R)xr - xi, b1 = xi - (R)xl;
  return a0 * b0 * fdl + a0 * b1 * fdr + a1 * b0 * ful + a1 * b1 * fur;
}

// Deformation in the value type R (FP64 plans, FP64 frames): ROWS pixels of
// one column per thread, in three phases -- every predictor load, then every
// frame tap, then the stores -- so a thread has ROWS independent load chains
// in flight instead of one.  Predictor by grid_bilinear, warp by warp_px
// (cuda_interpolate_frame's formula); masked pixels 0
// (cuda_interpolate_masked_frame, gpu_process.py:1461-1506).
template <typename R, typename T, int ROWS>
__global__ __launch_bounds__(256) void deform_rows_kernel(DeformArgsT<R> a) {
  // the block's patch of the predictor grid (rows and columns of the cells
  // its pixels fall in, plus one) staged in LDS once, when it fits
  constexpr int PR = 4, PS = 72;
  __shared__ R su[PR * PS], sv[PR * PS];
  const int Xb = blockIdx.x * blockDim.x;
  const int X = Xb + threadIdx.x;
  const int Y0 = blockIdx.y * ROWS;
  const R inv = R(1) / a.spacing;
  const int iy_lo = grid_cell((R)Y0, a.y0, inv, a.m);
  const int iy_hi = grid_cell((R)min(Y0 + ROWS - 1, a.ht - 1), a.y0, inv, a.m) + 1;
  const int ix_lo = grid_cell((R)Xb, a.x0, inv, a.n);
  const int ix_hi = grid_cell((R)min(Xb + (int)blockDim.x - 1, a.wd - 1), a.x0, inv, a.n) + 1;
  const bool staged = iy_hi - iy_lo < PR && ix_hi - ix_lo < PS;  // uniform over the block
  if (staged) {
    const int rows = iy_hi - iy_lo + 1, cols = ix_hi - ix_lo + 1;
    for (int e = threadIdx.x; e < rows * cols; e += blockDim.x) {
      const int r = e / cols, c = e - r * cols;
      const size_t g = (size_t)(iy_lo + r) * a.n + ix_lo + c;
      su[r * PS + c] = a.up[g];
      sv[r * PS + c] = a.vp[g];
    }
    __syncthreads();
  }
  if (X >= a.wd) return;
  R up[ROWS], vp[ROWS], oa[ROWS], ob[ROWS];
#pragma unroll
  for (int k = 0; k < ROWS; ++k) {
    const int Y = min(Y0 + k, a.ht - 1);
    if (staged)
      grid_bilinear_lds<R, PS>(su, sv, iy_lo, ix_lo, a.m, a.n, a.y0, a.x0, a.spacing, (R)Y, (R)X, up[k], vp[k]);
    else
      grid_bilinear<R>(a.up, a.vp, a.m, a.n, a.y0, a.x0, a.spacing, (R)Y, (R)X, up[k], vp[k]);
  }
#pragma unroll
  for (int k = 0; k < ROWS; ++k) {
    const int Y = min(Y0 + k, a.ht - 1);
    const size_t idx = (size_t)Y * a.wd + X;
    oa[k] = a.warp_a ? warp_px<R, T>(a.frame_a, a.ht, a.wd, (R)X, (R)Y, up[k], vp[k], a.dt_a)
                     : ldr<R, T>(a.frame_a, idx);
    ob[k] = a.warp_b ? warp_px<R, T>(a.frame_b, a.ht, a.wd, (R)X, (R)Y, up[k], vp[k], a.dt_b)
                     : ldr<R, T>(a.frame_b, idx);
  }
#pragma unroll
  for (int k = 0; k < ROWS; ++k) {
    const int Y = Y0 + k;
    if (Y >= a.ht) break;
    const size_t idx = (size_t)Y * a.wd + X;
    // cuda_interpolate_masked_frame (gpu_process.py:1461-1506) leaves masked
    // pixels at 0
    const bool masked = a.mask && a.mask[idx];
    a.out_a[idx] = masked ? R(0) : oa[k];
    a.out_b[idx] = masked ? R(0) : ob[k];
  }
}

// ---------------------------------------------------------------------------
// FP32 deformation of u8 / u16 / f32 frames, DR rows of one column per
// thread (the single-pair and the batched launch run this body).
//
// Predictor (RegularGridInterpolator semantics as grid_bilinear: clamped
// cell, unclamped fraction).  The thread's column fixes its grid cell ix and
// fraction tx once; the grid rows its DR rows fall between are interpolated
// along x once (h_r = f[g0 + r][ix] (1 - tx) + f[g0 + r][ix + 1] tx), and
// each row then needs (1 - ty) h_d + ty h_{d+1} with its block-uniform cell
// g0 + d and fraction ty.  Warp (cuda_interpolate_frame,
// gpu_process.py:1528-1546): clamped base pixel, unclamped fraction; the two
// horizontally adjacent taps of a row come from one buffer load (32-bit
// offsets, the lower row through the scalar offset), and the bilinear value
// is formed as a1 (b0 ful + b1 fur) + a0 (b0 fdl + b1 fdr).  Stores go
// through buffer stores with the row in the scalar offset.
// ---------------------------------------------------------------------------

// Two horizontally adjacent pixels starting at element idx (byte offset
// idx * sizeof(T)) of the row selected by the scalar byte offset soff: one
// load into a raw word, split into two floats afterwards.
template <typename T>
struct PixelPair;
template <>
struct PixelPair<uint8_t> {
  struct Raw {
    uint32_t a, b;
  };
  static __device__ __forceinline__ Raw load(rsrc_t r, uint32_t idx, uint32_t soff) {
    return Raw{__builtin_amdgcn_raw_buffer_load_b8(r, (int)idx, (int)soff, 0),
               __builtin_amdgcn_raw_buffer_load_b8(r, (int)idx + 1, (int)soff, 0)};
  }
  static __device__ __forceinline__ void split(Raw w, float& a, float& b) {
    a = (float)w.a;
    b = (float)w.b;
  }
};
template <>
struct PixelPair<uint16_t> {
  // two naturally aligned 2-byte loads: a 4-byte load at an odd pixel is
  // misaligned and costs more than the two
  struct Raw {
    uint32_t a, b;
  };
  static __device__ __forceinline__ Raw load(rsrc_t r, uint32_t idx, uint32_t soff) {
    return Raw{__builtin_amdgcn_raw_buffer_load_b16(r, (int)(idx * 2u), (int)soff, 0),
               __builtin_amdgcn_raw_buffer_load_b16(r, (int)(idx * 2u) + 2, (int)soff, 0)};
  }
  static __device__ __forceinline__ void split(Raw w, float& a, float& b) {
    a = (float)w.a;
    b = (float)w.b;
  }
};
template <>
struct PixelPair<float> {
  struct Raw {
    float a, b;
  };
  static __device__ __forceinline__ Raw load(rsrc_t r, uint32_t idx, uint32_t soff) {
    return Raw{buf_f32(r, idx * 4u, soff), buf_f32(r, idx * 4u + 4u, soff)};
  }
  static __device__ __forceinline__ void split(Raw w, float& a, float& b) {
    a = w.a;
    b = w.b;
  }
};

template <typename T>
struct DeformTile {
  static constexpr int DR = kDeformRows;  // rows per thread
  using PP = PixelPair<T>;

  // Predictor of the DR rows from y0 (RegularGridInterpolator semantics as
  // grid_bilinear): h_r = f[g0 + r][ix] (1 - tx) + f[g0 + r][ix + 1] tx for
  // the grid rows g0, g0 + 1, g0 + 2, then (1 - ty) h_d + ty h_{d+1} per row
  // (row cell g0 + d, block-uniform).  Lane k of the wave finds row k's cell
  // and fraction, read back as wave-uniform values.
  static __device__ __forceinline__ void predictor(const DeformArgs& a, int ix, float tx, float wx0, float inv, int y0,
                                                   float (&up)[DR], float (&vp)[DR]) {
    const rsrc_t ru = make_rsrc(a.up, (uint32_t)(a.m * a.n) * 4u);
    const rsrc_t rv = make_rsrc(a.vp, (uint32_t)(a.m * a.n) * 4u);
    const int g0 = __builtin_amdgcn_readfirstlane(grid_cell((float)y0, a.y0, inv, a.m));
    float hu[3], hv[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      const int gr = g0 + r < a.m ? g0 + r : a.m - 1;
      const uint32_t so = (uint32_t)(gr * a.n) * 4u;
      const float u0 = buf_f32(ru, (uint32_t)ix * 4u, so), u1 = buf_f32(ru, (uint32_t)ix * 4u + 4u, so);
      const float v0 = buf_f32(rv, (uint32_t)ix * 4u, so), v1 = buf_f32(rv, (uint32_t)ix * 4u + 4u, so);
      hu[r] = fmaf(u1, tx, u0 * wx0);
      hv[r] = fmaf(v1, tx, v0 * wx0);
    }
    const int lane = (int)(threadIdx.x & 63);
    const float Yl = (float)min(y0 + (lane < DR ? lane : 0), a.ht - 1);
    const int iyl = grid_cell(Yl, a.y0, inv, a.m);
    const float tyl = (Yl - (a.y0 + (float)iyl * a.spacing)) * inv;
#pragma unroll
    for (int k = 0; k < DR; ++k) {
      const int iy = __builtin_amdgcn_readlane(iyl, k);
      const float ty = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, tyl), k));
      const float wy0 = 1.0f - ty;
      const int d = iy - g0;  // block-uniform
      if (d == 0) {
        up[k] = fmaf(ty, hu[1], wy0 * hu[0]);
        vp[k] = fmaf(ty, hv[1], wy0 * hv[0]);
      } else if (d == 1) {
        up[k] = fmaf(ty, hu[2], wy0 * hu[1]);
        vp[k] = fmaf(ty, hv[2], wy0 * hv[1]);
      } else {  // grid spacing below DR pixels: the two grid rows directly
        const uint32_t s0 = (uint32_t)(iy * a.n) * 4u, s1 = s0 + (uint32_t)a.n * 4u;
        const float u00 = buf_f32(ru, (uint32_t)ix * 4u, s0), u01 = buf_f32(ru, (uint32_t)ix * 4u + 4u, s0);
        const float u10 = buf_f32(ru, (uint32_t)ix * 4u, s1), u11 = buf_f32(ru, (uint32_t)ix * 4u + 4u, s1);
        const float v00 = buf_f32(rv, (uint32_t)ix * 4u, s0), v01 = buf_f32(rv, (uint32_t)ix * 4u + 4u, s0);
        const float v10 = buf_f32(rv, (uint32_t)ix * 4u, s1), v11 = buf_f32(rv, (uint32_t)ix * 4u + 4u, s1);
        up[k] = fmaf(ty, fmaf(u11, tx, u10 * wx0), wy0 * fmaf(u01, tx, u00 * wx0));
        vp[k] = fmaf(ty, fmaf(v11, tx, v10 * wx0), wy0 * fmaf(v01, tx, v00 * wx0));
      }
    }
  }

  // One frame's taps at pixel (Xf, Yf) moved by (up, vp) * dt: the clamped
  // base pixel (floats hold these integers exactly; a NaN coordinate clamps
  // to 0 and yields a NaN value), the tap pairs of its two rows, the
  // unclamped fractions.  tap() only issues the loads; value() consumes them,
  // so that every row's loads are in flight before the first is used.
  struct Tap {
    typename PP::Raw top, bot;
    float b1, a0;
  };
  static __device__ __forceinline__ Tap tap(rsrc_t r, float ht2, float wd2, uint32_t wd, float Xf, float Yf, float up,
                                            float vp, float dt) {
    const float xi = fmaf(up, dt, Xf);
    const float yi = fmaf(vp, dt, Yf);
    const float xl = __builtin_amdgcn_fmed3f(floorf(xi), 0.0f, wd2);
    const float yu = __builtin_amdgcn_fmed3f(floorf(yi), 0.0f, ht2);
    const uint32_t idx = __umul24((uint32_t)yu, wd) + (uint32_t)xl;  // frames below 2^24 pixels a side
    Tap t;
    t.top = PP::load(r, idx, 0u);
    t.bot = PP::load(r, idx, wd * (uint32_t)sizeof(T));
    t.b1 = xi - xl;
    t.a0 = yi - yu;
    return t;
  }
  // (1 - a0) ((1 - b1) ful + b1 fur) + a0 ((1 - b1) fdl + b1 fdr), as two
  // horizontal lerps and a vertical one
  static __device__ __forceinline__ float value(const Tap& t) {
    float ful, fur, fdl, fdr;
    PP::split(t.top, ful, fur);
    PP::split(t.bot, fdl, fdr);
    const float top = fmaf(t.b1, fur - ful, ful);
    const float bot = fmaf(t.b1, fdr - fdl, fdl);
    return fmaf(t.a0, bot - top, top);
  }

  // the unwarped pixel X of the row starting at element row_off
  static __device__ __forceinline__ float raw(rsrc_t r, uint32_t X, uint32_t row_off) {
    if constexpr (sizeof(T) == 4)
      return buf_f32(r, X * 4u, row_off * 4u);
    else if constexpr (sizeof(T) == 2)
      return (float)(uint32_t)__builtin_amdgcn_raw_buffer_load_b16(r, (int)(X * 2u), (int)(row_off * 2u), 0);
    else
      return (float)(uint32_t)__builtin_amdgcn_raw_buffer_load_b8(r, (int)X, (int)row_off, 0);
  }

  // Rows by * DR .. by * DR + DR - 1 of column bx * blockDim.x + tid; stores
  // leave masked pixels at 0 (cuda_interpolate_masked_frame,
  // gpu_process.py:1461-1506).
  static __device__ void run(const DeformArgs& a, int bx, int by) {
    const int X = bx * (int)blockDim.x + (int)threadIdx.x;
    // every lane of the wave forms the predictor (lane k supplies row k's
    // cell and fraction to the others); columns beyond the frame leave after
    if (X - (int)(threadIdx.x & 63) >= a.wd) return;  // the whole wave is beyond the frame
    const int Y0 = by * DR;
    const float inv = 1.0f / a.spacing;
    const float Xf = (float)X;
    const int ix = grid_cell(Xf, a.x0, inv, a.n);
    const float tx = (Xf - (a.x0 + (float)ix * a.spacing)) * inv;
    const float wx0 = 1.0f - tx;
    float up[DR], vp[DR];
    predictor(a, ix, tx, wx0, inv, Y0, up, vp);
    if (X >= a.wd) return;  // no barrier below

    const uint32_t npx = (uint32_t)a.ht * (uint32_t)a.wd;
    const rsrc_t fa = make_rsrc(a.frame_a, npx * (uint32_t)sizeof(T));
    const rsrc_t fb = make_rsrc(a.frame_b, npx * (uint32_t)sizeof(T));
    const float ht2 = (float)(a.ht - 2), wd2 = (float)(a.wd - 2);
    float oa[DR], ob[DR];
    if (a.warp_a && a.warp_b) {  // both frames move (0 < deforming_par < 1): every load first
      Tap ta[DR], tb[DR];
#pragma unroll
      for (int k = 0; k < DR; ++k) {
        const float Yf = (float)min(Y0 + k, a.ht - 1);
        ta[k] = tap(fa, ht2, wd2, (uint32_t)a.wd, Xf, Yf, up[k], vp[k], a.dt_a);
        tb[k] = tap(fb, ht2, wd2, (uint32_t)a.wd, Xf, Yf, up[k], vp[k], a.dt_b);
      }
#pragma unroll
      for (int k = 0; k < DR; ++k) {
        oa[k] = value(ta[k]);
        ob[k] = value(tb[k]);
      }
    } else {  // deforming_par 0 or 1: one frame copied unchanged
#pragma unroll
      for (int k = 0; k < DR; ++k) {
        const int Y = min(Y0 + k, a.ht - 1);
        const float Yf = (float)Y;
        const uint32_t row = (uint32_t)Y * (uint32_t)a.wd;
        oa[k] = a.warp_a ? value(tap(fa, ht2, wd2, (uint32_t)a.wd, Xf, Yf, up[k], vp[k], a.dt_a))
                         : raw(fa, (uint32_t)X, row);
        ob[k] = a.warp_b ? value(tap(fb, ht2, wd2, (uint32_t)a.wd, Xf, Yf, up[k], vp[k], a.dt_b))
                         : raw(fb, (uint32_t)X, row);
      }
    }
    const rsrc_t oa_r = make_rsrc(a.out_a, npx * 4u);
    const rsrc_t ob_r = make_rsrc(a.out_b, npx * 4u);
    const rsrc_t mk = make_rsrc(a.mask, a.mask ? npx : 0u);
#pragma unroll
    for (int k = 0; k < DR; ++k) {
      const int Y = Y0 + k;
      if (Y >= a.ht) break;
      const uint32_t row = (uint32_t)Y * (uint32_t)a.wd;
      float va = oa[k], vb = ob[k];
      if (a.mask) {
        const uint32_t mv = __builtin_amdgcn_raw_buffer_load_b8(mk, X, (int)row, 0);
        if (mv) va = vb = 0.0f;
      }
      __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, va), oa_r, X * 4, (int)(row * 4u), 0);
      __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, vb), ob_r, X * 4, (int)(row * 4u), 0);
    }
  }
};

template <typename T>
__global__ __launch_bounds__(256) void deform_tile_kernel(DeformArgs a) {
  DeformTile<T>::run(a, blockIdx.x, blockIdx.y);
}

// Many frames in one launch: grid.y runs over the concatenated row groups
// (kDeformRows rows each) of the frames; TW: the workgroup width.
template <typename T, int TW>
__global__ __launch_bounds__(TW) void deform_batch_kernel(const DeformArgs* fields, const int* group_prefix,
                                                          int n_fields) {
  const int r = blockIdx.y;
  int lo = 0, hi = n_fields - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (group_prefix[mid] <= r) lo = mid;
    else hi = mid - 1;
  }
  const DeformArgs a = fields[lo];
  DeformTile<T>::run(a, blockIdx.x, r - group_prefix[lo]);
}

template <typename R>
struct GridInterpBody {
  using Args = GridInterpArgsT<R>;
  static __device__ void run(const Args& a, VB b) {
    const int k = b.bx * blockDim.x + threadIdx.x;
    if (k >= a.m_out * a.n_out) return;
    const int i = k / a.n_out, j = k % a.n_out;
    const R yq = a.y0_out + (R)i * a.sp_out;
    const R xq = a.x0_out + (R)j * a.sp_out;
    R ou, ov;
    grid_bilinear<R>(a.u_in, a.v_in, a.m_in, a.n_in, a.y0_in, a.x0_in, a.sp_in, yq, xq, ou, ov);
    a.u_out[k] = ou;
    a.v_out[k] = ov;
  }
};
template <typename R>
__global__ __launch_bounds__(256) void grid_interp_kernel(GridInterpArgsT<R> a) {
  GridInterpBody<R>::run(a, vb_native());
}

// Median of the valid entries of a[0..cnt) after sorting (cnt <= 8).
template <typename R>
__device__ __forceinline__ void cswap(R& x, R& y) {
  const R lo = vmin(x, y), hi = vmax(x, y);
  x = lo;
  y = hi;
}
template <typename R>
__device__ __forceinline__ R median8(R (&s)[8], int cnt) {
  // Batcher odd-even merge sort network for 8 keys (19 comparators); missing
  // entries are +inf and sort to the end.
  cswap(s[0], s[1]); cswap(s[2], s[3]); cswap(s[4], s[5]); cswap(s[6], s[7]);
  cswap(s[0], s[2]); cswap(s[1], s[3]); cswap(s[4], s[6]); cswap(s[5], s[7]);
  cswap(s[1], s[2]); cswap(s[5], s[6]);
  cswap(s[0], s[4]); cswap(s[1], s[5]); cswap(s[2], s[6]); cswap(s[3], s[7]);
  cswap(s[2], s[4]); cswap(s[3], s[5]);
  cswap(s[1], s[2]); cswap(s[3], s[4]); cswap(s[5], s[6]);
  if (cnt <= 0) return vnan<R>();
  R r = R(0);
  const R h = R(0.5);
  switch (cnt) {
    case 1: r = s[0]; break;
    case 2: r = h * (s[0] + s[1]); break;
    case 3: r = s[1]; break;
    case 4: r = h * (s[1] + s[2]); break;
    case 5: r = s[2]; break;
    case 6: r = h * (s[2] + s[3]); break;
    case 7: r = s[3]; break;
    default: r = h * (s[3] + s[4]); break;
  }
  return r;
}

// ValidationGPU.median_validation (gpu_validation.py:136-141, 166-205):
// NaN-padded 3x3 neighbourhood, centre excluded, nanmedian, absolute test,
// flags OR-ed over u and v.
__global__ __launch_bounds__(256) void median_kernel(const float* u, const float* v, int m, int n, float tol,
                                                     uint8_t* flags, int32_t* count) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  bool flag = false;
  if (k < m * n) {
    const int i = k / n, j = k % n;
    float su[8], sv[8];
    int cu = 0, cv = 0;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      su[q] = INFINITY;
      sv[q] = INFINITY;
    }
#pragma unroll
    for (int di = -1; di <= 1; ++di) {
#pragma unroll
      for (int dj = -1; dj <= 1; ++dj) {
        if (di == 0 && dj == 0) continue;
        const int ii = i + di, jj = j + dj;
        if (ii < 0 || ii >= m || jj < 0 || jj >= n) continue;
        const float a = u[ii * n + jj], b = v[ii * n + jj];
        if (!isnan(a)) {
          // insert at position cu (register array indexed by a small runtime
          // value -> keep it branch-free with a select chain)
#pragma unroll
          for (int q = 0; q < 8; ++q) su[q] = (q == cu) ? a : su[q];
          ++cu;
        }
        if (!isnan(b)) {
#pragma unroll
          for (int q = 0; q < 8; ++q) sv[q] = (q == cv) ? b : sv[q];
          ++cv;
        }
      }
    }
    const float mu = median8(su, cu);
    const float mv = median8(sv, cv);
    flag = (fabsf(u[k] - mu) > tol) || (fabsf(v[k] - mv) > tol);
    flags[k] = flag ? 1 : 0;
  }
  const unsigned long long b = __ballot(flag);
  if ((threadIdx.x & 63) == 0 && b) atomicAdd(count, (int32_t)__popcll(b));
}

// ---------------------------------------------------------------------------
// Spring replacement (ReplacementGPU.spring_replacement,
// gpu_validation.py:336-371, 381-415; PIVGPU.replace_outliers,
// gpu_process.py:788-834).  The reference links each outlier k to the four
// vectors N4(k) that lie s = replacing_size rows or columns away
// (:279-283) and assembles
//     x_k - (1/p_k) sum_{l in N4(k), l outlier} x_l = rhs_k,
// rhs_k the NaN-mean over the in-bounds N4(k) with the outliers set to 0,
// and 1/p_k by position (:365-371): p_k = 2 in the four s x s corner
// blocks, 3 in the rest of the s-wide border, 4 inside -- the in-bounds
// link count whenever the field is at least 2 s wide, at least that count
// otherwise.  Scaling row k by p_k gives a symmetric matrix, diagonally
// dominant, positive definite whenever each linked outlier cluster touches a
// valid vector or the border, solved here by Jacobi-preconditioned CG in
// FP64 for u and v together, in one workgroup.  When every vector is an
// outlier the system is singular; the reference's NaN -> reset path then
// leaves the field unchanged, and so does this kernel.
// ---------------------------------------------------------------------------
constexpr int kSpringLds = 1024;  // outlier count handled by the LDS-resident solver
constexpr int kClusterMax = 6;    // largest outlier cluster solved directly (spring_clusters)

template <int K>
__device__ __forceinline__ void block_sum(double (&v)[K], double (*scratch)[K]) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nw = blockDim.x >> 6;
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = wave_sum(v[k]);
  __syncthreads();
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) scratch[wave][k] = v[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) {
    double s = 0.0;
    for (int w = 0; w < nw; ++w) s += scratch[w][k];
    v[k] = s;
  }
}

// CG over global-memory vectors for large outlier sets, run by the 256-thread
// workgroup after the compaction (slot map and list already written).
template <int K>
__device__ __forceinline__ void sum256(double (&v)[K], double (*scratch)[4]) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = wave_sum(v[k]);
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) scratch[wave][k] = v[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = scratch[0][k] + scratch[1][k] + scratch[2][k] + scratch[3][k];
  __syncthreads();
}

// Outlier tests: the slot map of the global-memory solver, or the flags
// themselves (a vector is an outlier exactly when it is flagged).
struct SlotOutlier {
  const int32_t* slot;
  __device__ bool operator()(int l) const { return slot[l] >= 0; }
};
struct FlagOutlier {
  const uint8_t* flags;
  __device__ bool operator()(int l) const { return flags[l] != 0; }
};

// Diagonal p_k of the scaled spring row at (i, j) for link distance s
// (gpu_validation.py:365-371).
__device__ __forceinline__ int spring_diag(int i, int j, int m, int n, int s) {
  const bool bi = i < s || i >= m - s, bj = j < s || j >= n - s;
  return (bi && bj) ? 2 : ((bi || bj) ? 3 : 4);
}

// Right-hand side of outlier row k scaled by p_k (spring_diag):
// (p_k / c_k) times the sum of the neighbours that are
// not outliers and not NaN, c_k counting the in-bounds neighbours that are
// outliers or not NaN -- the reference's NaN-mean over the neighbours with
// the outliers set to 0 (gpu_validation.py:342, 381-394).  c_k = 0 gives a
// NaN right-hand side there; such a row has no outlier neighbour, so it is
// decoupled, and the vector keeps its value (the reset of unresolved
// vectors, gpu_process.py:829-832): its right-hand side is 0 here and
// nan_isolated() tells the final write to skip it.
template <typename R, typename Out>
__device__ __forceinline__ void spring_rhs(const R* u, const R* v, Out is_out, int m, int n, int pk,
                                           const int (&ni)[4], const int (&nj)[4], double* bu, double* bv,
                                           bool* iso_u = nullptr, bool* iso_v = nullptr) {
  // every load is issued up front (out-of-bounds neighbours read a clamped,
  // unused element), then the contributions are masked
  bool inb[4], out[4];
  R a[4], b[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    inb[e] = ni[e] >= 0 && ni[e] < m && nj[e] >= 0 && nj[e] < n;
    const int l = min(max(ni[e], 0), m - 1) * n + min(max(nj[e], 0), n - 1);
    out[e] = is_out(l);
    a[e] = u[l];
    b[e] = v[l];
  }
  double su = 0.0, sv = 0.0;
  int cnt = 0, cu = 0, cv = 0;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    if (!inb[e]) continue;
    ++cnt;
    if (out[e]) {
      ++cu;
      ++cv;
    } else {
      if (!isnan(a[e])) {
        su += (double)a[e];
        ++cu;
      }
      if (!isnan(b[e])) {
        sv += (double)b[e];
        ++cv;
      }
    }
  }
  (void)cnt;
  *bu = (cu > 0 && cu != pk) ? su * pk / cu : su;
  *bv = (cv > 0 && cv != pk) ? sv * pk / cv : sv;
  if (iso_u) *iso_u = cu == 0;  // what nan_isolated() finds
  if (iso_v) *iso_v = cv == 0;
}

template <typename R, typename Out>
__device__ __forceinline__ void nan_isolated(const R* u, const R* v, Out is_out, int m, int n, int sz, int k,
                                             bool* ku, bool* kv) {
  const int i = k / n, j = k % n;
  const int ni[4] = {i, i, i + sz, i - sz};
  const int nj[4] = {j - sz, j + sz, j, j};
  int cu = 0, cv = 0;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    if (ni[e] < 0 || ni[e] >= m || nj[e] < 0 || nj[e] >= n) continue;
    const int l = ni[e] * n + nj[e];
    if (is_out(l)) {
      ++cu;
      ++cv;
    } else {
      cu += isnan(u[l]) ? 0 : 1;
      cv += isnan(v[l]) ? 0 : 1;
    }
  }
  *ku = cu == 0;
  *kv = cv == 0;
}

template <typename R>
__device__ void spring_cg_global(R* u, R* v, int m, int n, int sz, int nv, const ReplaceWork& w,
                                 double (*red)[4], const int32_t* cls) {
  constexpr int T = 256;
  const int tid = threadIdx.x;
  double* xu = w.work;
  double* xv = w.work + (size_t)nv;
  double* ru = w.work + 2 * (size_t)nv;
  double* rv = w.work + 3 * (size_t)nv;
  double* pu = w.work + 4 * (size_t)nv;
  double* pv = w.work + 5 * (size_t)nv;
  double* qu = w.work + 6 * (size_t)nv;
  double* qv = w.work + 7 * (size_t)nv;
  auto nbrs = [&](int s, int (&ni)[4], int (&nj)[4]) {
    const int k = w.list[s];
    const int i = k / n, j = k % n;
    ni[0] = i; nj[0] = j - sz;
    ni[1] = i; nj[1] = j + sz;
    ni[2] = i + sz; nj[2] = j;
    ni[3] = i - sz; nj[3] = j;
  };
  double acc[4] = {0, 0, 0, 0};
  for (int s = tid; s < nv; s += T) {
    int ni[4], nj[4];
    nbrs(s, ni, nj);
    double bu = 0.0, bv = 0.0;
    const int kk = w.list[s];
    const int nk = spring_diag(kk / n, kk % n, m, n, sz);
    spring_rhs(u, v, SlotOutlier{w.slot}, m, n, nk, ni, nj, &bu, &bv);
    xu[s] = 0.0;
    xv[s] = 0.0;
    ru[s] = bu;
    rv[s] = bv;
    const double zu = bu / nk, zv = bv / nk;
    pu[s] = zu;
    pv[s] = zv;
    acc[0] += bu * zu;
    acc[1] += bv * zv;
    acc[2] += bu * bu;
    acc[3] += bv * bv;
  }
  sum256<4>(acc, red);
  double rz_u = acc[0], rz_v = acc[1];
  const double tol2_u = 1e-20 * acc[2], tol2_v = 1e-20 * acc[3];
  bool done_u = acc[2] == 0.0, done_v = acc[3] == 0.0;
  const int max_it = nv + 200 < 20000 ? nv + 200 : 20000;
  for (int it = 0; it < max_it && !(done_u && done_v); ++it) {
    double pq[2] = {0, 0};
    for (int s = tid; s < nv; s += T) {
      int ni[4], nj[4];
      nbrs(s, ni, nj);
      const int kk = w.list[s];
      const int nk = spring_diag(kk / n, kk % n, m, n, sz);
      double au = 0.0, av = 0.0;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        if (ni[e] < 0 || ni[e] >= m || nj[e] < 0 || nj[e] >= n) continue;
        const int sl = w.slot[ni[e] * n + nj[e]];
        if (sl >= 0) {
          au += pu[sl];
          av += pv[sl];
        }
      }
      const double qa = nk * pu[s] - au, qb = nk * pv[s] - av;
      qu[s] = qa;
      qv[s] = qb;
      pq[0] += pu[s] * qa;
      pq[1] += pv[s] * qb;
    }
    double pq4[4] = {pq[0], pq[1], 0.0, 0.0};
    sum256<4>(pq4, red);
    const double alpha_u = (done_u || pq4[0] == 0.0) ? 0.0 : rz_u / pq4[0];
    const double alpha_v = (done_v || pq4[1] == 0.0) ? 0.0 : rz_v / pq4[1];
    double nr[4] = {0, 0, 0, 0};
    for (int s = tid; s < nv; s += T) {
      const int k = w.list[s];
      const int nk = spring_diag(k / n, k % n, m, n, sz);
      xu[s] += alpha_u * pu[s];
      xv[s] += alpha_v * pv[s];
      const double r1 = ru[s] - alpha_u * qu[s], r2 = rv[s] - alpha_v * qv[s];
      ru[s] = r1;
      rv[s] = r2;
      nr[0] += r1 * r1 / nk;
      nr[1] += r2 * r2 / nk;
      nr[2] += r1 * r1;
      nr[3] += r2 * r2;
    }
    sum256<4>(nr, red);
    const double beta_u = rz_u > 0.0 ? nr[0] / rz_u : 0.0;
    const double beta_v = rz_v > 0.0 ? nr[1] / rz_v : 0.0;
    rz_u = nr[0];
    rz_v = nr[1];
    done_u = done_u || nr[2] <= tol2_u;
    done_v = done_v || nr[3] <= tol2_v;
    for (int s = tid; s < nv; s += T) {
      const int k = w.list[s];
      const int nk = spring_diag(k / n, k % n, m, n, sz);
      pu[s] = ru[s] / nk + beta_u * pu[s];
      pv[s] = rv[s] / nk + beta_v * pv[s];
    }
    __syncthreads();
  }
  __syncthreads();
  for (int s = tid; s < nv; s += T) {
    const int k = w.list[s];
    bool ku, kv;
    nan_isolated(u, v, SlotOutlier{w.slot}, m, n, sz, k, &ku, &kv);
    if (cls && cls[(k / n) % sz * sz + (k % n) % sz]) continue;  // singular block: keeps its values
    if (!ku) u[k] = (R)xu[s];
    if (!kv) v[k] = (R)xv[s];
  }
}

// LDS-resident variant of the same solve for up to kSpringLds outliers (the
// common case): all CG vectors and the outlier neighbour lists live in LDS, so
// an iteration touches no global memory.  The flags are compacted by the whole
// 256-thread workgroup with 16-byte loads into a row-major outlier list in
// LDS (no slot map: a neighbour's slot is found by binary search in the
// list); up to 256 outliers are then solved by wave 0 alone (shuffle
// reductions, no workgroup barriers).  CG starts from x0 = D^-1 b (the mean
// of the valid neighbours), exact for isolated outliers.
constexpr int kSpringLdsThreads = 256;

struct SpringLds {
  // xu, xv, pu, pv carry one more element, held at 0: neighbour gathers read
  // it for neighbours that are not outliers
  double xu[kSpringLds + 1], xv[kSpringLds + 1], ru[kSpringLds], rv[kSpringLds];
  double pu[kSpringLds + 1], pv[kSpringLds + 1], qu[kSpringLds], qv[kSpringLds];
  int16_t nb[kSpringLds][4];
  uint8_t nk[kSpringLds];
  uint8_t iso[kSpringLds];  // bit 0: u keeps its value, bit 1: v does
  int32_t list[kSpringLds];
  int32_t scan[kSpringLdsThreads / 64];
  int32_t cls[kMaxNeighbourhoodSize * kMaxNeighbourhoodSize];  // singular residue classes
  double red[4][4];
  // cluster solve of up to 256 outliers (spring_clusters)
  int16_t lab[256];        // cluster label: smallest member slot
  int32_t ccount[256];     // members per label
  int16_t cmem[256][kClusterMax];  // members per label (the first kClusterMax)
  int16_t croot[256];      // labels of the clusters solved directly
};

__device__ __forceinline__ void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

template <int K, int T>
__device__ __forceinline__ void group_sum(double (&v)[K], double (*scratch)[4]) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = wave_sum(v[k]);
  if constexpr (T > 64) {
    if (lane == 0) {
#pragma unroll
      for (int k = 0; k < K; ++k) scratch[wave][k] = v[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = scratch[0][k] + scratch[1][k] + scratch[2][k] + scratch[3][k];
    __syncthreads();
  }
}

// CG iterations over the LDS system with T active threads (64: wave 0 only).
template <int T>
__device__ void spring_cg(SpringLds& L, int nv) {
  const int tid = threadIdx.x;
  auto sync = [] {
    if constexpr (T > 64) __syncthreads(); else wave_sync_lds();
  };
  double acc[4] = {0, 0, 0, 0};
  for (int s = tid; s < nv; s += T) {
    double au = 0.0, av = 0.0;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int c = L.nb[s][e];
      if (c >= 0) {
        au += L.xu[c];
        av += L.xv[c];
      }
    }
    L.ru[s] = au;  // r0 = b - A x0 = sum of x0 over outlier neighbours
    L.rv[s] = av;
    const double inv = 1.0 / L.nk[s];
    L.pu[s] = au * inv;
    L.pv[s] = av * inv;
    acc[0] += au * au * inv;
    acc[1] += av * av * inv;
    const double bu = L.xu[s] * L.nk[s], bv = L.xv[s] * L.nk[s];  // |b|^2 as the scale
    acc[2] += bu * bu;
    acc[3] += bv * bv;
  }
  group_sum<4, T>(acc, L.red);
  double rz_u = acc[0], rz_v = acc[1];
  const double tol2_u = 1e-20 * acc[2], tol2_v = 1e-20 * acc[3];
  bool done_u = rz_u <= tol2_u, done_v = rz_v <= tol2_v;
  const int max_it = nv + 200;
  for (int it = 0; it < max_it && !(done_u && done_v); ++it) {
    sync();
    double pq[2] = {0, 0};
    for (int s = tid; s < nv; s += T) {
      double au = 0.0, av = 0.0;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int c = L.nb[s][e];
        if (c >= 0) {
          au += L.pu[c];
          av += L.pv[c];
        }
      }
      const double qa = L.nk[s] * L.pu[s] - au, qb = L.nk[s] * L.pv[s] - av;
      L.qu[s] = qa;
      L.qv[s] = qb;
      pq[0] += L.pu[s] * qa;
      pq[1] += L.pv[s] * qb;
    }
    group_sum<2, T>(pq, L.red);
    const double alpha_u = (done_u || pq[0] == 0.0) ? 0.0 : rz_u / pq[0];
    const double alpha_v = (done_v || pq[1] == 0.0) ? 0.0 : rz_v / pq[1];
    double nr[4] = {0, 0, 0, 0};
    for (int s = tid; s < nv; s += T) {
      L.xu[s] += alpha_u * L.pu[s];
      L.xv[s] += alpha_v * L.pv[s];
      const double r1 = L.ru[s] - alpha_u * L.qu[s], r2 = L.rv[s] - alpha_v * L.qv[s];
      L.ru[s] = r1;
      L.rv[s] = r2;
      const double inv = 1.0 / L.nk[s];
      nr[0] += r1 * r1 * inv;
      nr[1] += r2 * r2 * inv;
      nr[2] += r1 * r1;
      nr[3] += r2 * r2;
    }
    group_sum<4, T>(nr, L.red);
    const double beta_u = rz_u > 0.0 ? nr[0] / rz_u : 0.0;
    const double beta_v = rz_v > 0.0 ? nr[1] / rz_v : 0.0;
    rz_u = nr[0];
    rz_v = nr[1];
    done_u = done_u || nr[2] <= tol2_u;
    done_v = done_v || nr[3] <= tol2_v;
    for (int s = tid; s < nv; s += T) {
      const double inv = 1.0 / L.nk[s];
      L.pu[s] = L.ru[s] * inv + beta_u * L.pu[s];
      L.pv[s] = L.rv[s] * inv + beta_v * L.pv[s];
    }
  }
  sync();
}

// Outlier clusters (vectors joined by outlier links) of 2..kClusterMax
// vectors, solved exactly on wave 0 before the CG iterations: the system is
// block diagonal by cluster, so CG started from these solutions converges at
// once unless a larger cluster remains (its rows start from x = b / n_k as
// before).  Clusters are labelled by their smallest slot (minimum-label
// propagation with pointer jumping), their members listed per label, and
// each lane takes one cluster: members sorted by slot (so the system and its
// rounding do not depend on the listing order), Gaussian elimination without
// pivoting (symmetric, diagonally dominant), every member's value written.
// Stores b = n_k x in ru / rv first (the right-hand sides the CG uses).
// Returns whether a larger cluster is left (then CG runs; otherwise every
// cluster is solved: single vectors by x = b / n_k already).

__device__ __noinline__ bool spring_clusters(SpringLds& L, int nv) {
  constexpr int E = 256 / 64, C = kClusterMax;
  const int lane = threadIdx.x;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int s = lane + 64 * e;
    if (s < nv) {
      const double nk = (double)L.nk[s];
      L.ru[s] = L.xu[s] * nk;
      L.rv[s] = L.xv[s] * nk;
      L.lab[s] = (int16_t)s;
      L.ccount[s] = 0;
    }
  }
  wave_sync_lds();
  // labels: the smallest slot of each cluster (a label only decreases; stops
  // when no lane changes one, after at most nv rounds)
  for (int it = 0; it < nv; ++it) {
    bool changed = false;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int s = lane + 64 * e;
      if (s < nv) {
        int m = L.lab[s];
#pragma unroll
        for (int d = 0; d < 4; ++d) {
          const int c = L.nb[s][d];
          if (c >= 0) m = min(m, (int)L.lab[c]);
        }
        m = min(m, (int)L.lab[m]);
        if (m < L.lab[s]) {
          L.lab[s] = (int16_t)m;
          changed = true;
        }
      }
    }
    wave_sync_lds();
    if (!__any(changed)) break;
  }
  // members per label (listing order is free: the solver sorts them)
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int s = lane + 64 * e;
    if (s < nv) {
      const int root = L.lab[s];
      const int k = atomicAdd(&L.ccount[root], 1);
      if (k < C) L.cmem[root][k] = (int16_t)s;
    }
  }
  wave_sync_lds();
  // clusters of 2..C vectors, one per lane
  int nroots = 0;
  bool larger = false;  // a cluster of more than C vectors is left to CG
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int s = lane + 64 * e;
    larger = larger || (s < nv && L.lab[s] == s && L.ccount[s] > C);
    const bool take = s < nv && L.lab[s] == s && L.ccount[s] >= 2 && L.ccount[s] <= C;
    const uint64_t bal = __ballot(take);
    if (take) L.croot[nroots + __popcll(bal & ((1ull << lane) - 1))] = (int16_t)s;
    nroots += __popcll(bal);
  }
  wave_sync_lds();
  for (int base = 0; base < nroots; base += 64) {
    const int ci = base + lane;
    if (ci < nroots) {
      const int root = L.croot[ci];
      const int cnt = L.ccount[root];
      int q[C];
#pragma unroll
      for (int j = 0; j < C; ++j) q[j] = j < cnt ? (int)L.cmem[root][j] : INT_MAX;
#pragma unroll
      for (int r = 0; r < C; ++r) {  // odd-even transposition sort
#pragma unroll
        for (int j = r & 1; j + 1 < C; j += 2) {
          const int lo = min(q[j], q[j + 1]), hi = max(q[j], q[j + 1]);
          q[j] = lo;
          q[j + 1] = hi;
        }
      }
      double A[C][C], bu[C], bv[C];
#pragma unroll
      for (int p = 0; p < C; ++p) {
        const bool real = p < cnt;
        const int sp = real ? q[p] : 0;
#pragma unroll
        for (int j = 0; j < C; ++j) A[p][j] = 0.0;
        A[p][p] = real ? (double)L.nk[sp] : 1.0;
        bu[p] = real ? L.ru[sp] : 0.0;
        bv[p] = real ? L.rv[sp] : 0.0;
#pragma unroll
        for (int d = 0; d < 4; ++d) {
          const int c = real ? (int)L.nb[sp][d] : -1;
#pragma unroll
          for (int j = 0; j < C; ++j)
            if (c >= 0 && q[j] == c) A[p][j] -= 1.0;
        }
      }
      double piv[C];
#pragma unroll
      for (int k = 0; k < C; ++k) {
        piv[k] = 1.0 / A[k][k];
#pragma unroll
        for (int i = k + 1; i < C; ++i) {
          const double f = A[i][k] * piv[k];
#pragma unroll
          for (int j = k + 1; j < C; ++j) A[i][j] = fma(-f, A[k][j], A[i][j]);
          bu[i] = fma(-f, bu[k], bu[i]);
          bv[i] = fma(-f, bv[k], bv[i]);
        }
      }
      double xu[C], xv[C];
#pragma unroll
      for (int k = C - 1; k >= 0; --k) {
        double su = bu[k], sv = bv[k];
#pragma unroll
        for (int j = k + 1; j < C; ++j) {
          su = fma(-A[k][j], xu[j], su);
          sv = fma(-A[k][j], xv[j], sv);
        }
        xu[k] = su * piv[k];
        xv[k] = sv * piv[k];
      }
#pragma unroll
      for (int p = 0; p < C; ++p) {
        if (p < cnt) {
          L.xu[q[p]] = xu[p];
          L.xv[q[p]] = xv[p];
        }
      }
    }
  }
  wave_sync_lds();
  return __any(larger);
}

// The same iterations for up to 256 outliers on wave 0 alone, with every
// per-outlier value in registers (lane l owns outliers l, l + 64, ...); only
// the search direction goes through LDS for the neighbour gathers.  Same
// arithmetic and summation order as spring_cg<64>; 1 / n_k from a table
// (the correctly rounded reciprocals of 1..4).
__device__ void spring_cg_wave(SpringLds& L, int nv) {
  constexpr int E = 256 / 64;
  const int lane = threadIdx.x;
  double xu[E], xv[E], ru[E], rv[E], pu[E], pv[E], nk[E], inv[E];
  int nb[E][4];  // gather index: the neighbour's slot, or the zero element
  bool act[E];
  double acc[4] = {0, 0, 0, 0};
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int s = lane + 64 * e;
    act[e] = s < nv;
    const int sc = act[e] ? s : 0;
    const int cnt = L.nk[sc];
    nk[e] = (double)cnt;
    inv[e] = cnt == 4 ? 0.25 : (cnt == 3 ? 1.0 / 3.0 : (cnt == 2 ? 0.5 : 1.0));
    xu[e] = L.xu[sc];
    xv[e] = L.xv[sc];
#pragma unroll
    for (int d = 0; d < 4; ++d) {
      const int c = L.nb[sc][d];
      nb[e][d] = c >= 0 ? c : kSpringLds;
    }
  }
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int sc = act[e] ? lane + 64 * e : 0;
    double au = 0.0, av = 0.0;
#pragma unroll
    for (int d = 0; d < 4; ++d) {  // + 0.0 for a neighbour that is no outlier
      au += L.xu[nb[e][d]];
      av += L.xv[nb[e][d]];
    }
    // r0 = b - A x = b - n_k x + (sum of x over outlier neighbours); exactly
    // the neighbour sum where x is still b / n_k
    const double bu = L.ru[sc], bv = L.rv[sc];
    const double r_u = (bu - nk[e] * xu[e]) + au, r_v = (bv - nk[e] * xv[e]) + av;
    ru[e] = r_u;
    rv[e] = r_v;
    pu[e] = r_u * inv[e];
    pv[e] = r_v * inv[e];
    if (act[e]) {
      acc[0] += r_u * r_u * inv[e];
      acc[1] += r_v * r_v * inv[e];
      acc[2] += bu * bu;  // |b|^2 as the scale
      acc[3] += bv * bv;
    }
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) acc[k] = wave_sum(acc[k]);
  double rz_u = acc[0], rz_v = acc[1];
  const double tol2_u = 1e-20 * acc[2], tol2_v = 1e-20 * acc[3];
  bool done_u = rz_u <= tol2_u, done_v = rz_v <= tol2_v;
  const int max_it = nv + 200;
  for (int it = 0; it < max_it && !(done_u && done_v); ++it) {
    wave_sync_lds();  // earlier gathers done before p is overwritten
#pragma unroll
    for (int e = 0; e < E; ++e) {
      if (act[e]) {
        L.pu[lane + 64 * e] = pu[e];
        L.pv[lane + 64 * e] = pv[e];
      }
    }
    wave_sync_lds();
    double pq[2] = {0, 0}, qu[E], qv[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
      double au = 0.0, av = 0.0;
#pragma unroll
      for (int d = 0; d < 4; ++d) {
        au += L.pu[nb[e][d]];
        av += L.pv[nb[e][d]];
      }
      qu[e] = nk[e] * pu[e] - au;
      qv[e] = nk[e] * pv[e] - av;
      if (act[e]) {
        pq[0] += pu[e] * qu[e];
        pq[1] += pv[e] * qv[e];
      }
    }
    pq[0] = wave_sum(pq[0]);
    pq[1] = wave_sum(pq[1]);
    const double alpha_u = (done_u || pq[0] == 0.0) ? 0.0 : rz_u / pq[0];
    const double alpha_v = (done_v || pq[1] == 0.0) ? 0.0 : rz_v / pq[1];
    double nr[4] = {0, 0, 0, 0};
#pragma unroll
    for (int e = 0; e < E; ++e) {
      xu[e] += alpha_u * pu[e];
      xv[e] += alpha_v * pv[e];
      ru[e] = ru[e] - alpha_u * qu[e];
      rv[e] = rv[e] - alpha_v * qv[e];
      if (act[e]) {
        nr[0] += ru[e] * ru[e] * inv[e];
        nr[1] += rv[e] * rv[e] * inv[e];
        nr[2] += ru[e] * ru[e];
        nr[3] += rv[e] * rv[e];
      }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) nr[k] = wave_sum(nr[k]);
    const double beta_u = rz_u > 0.0 ? nr[0] / rz_u : 0.0;
    const double beta_v = rz_v > 0.0 ? nr[1] / rz_v : 0.0;
    rz_u = nr[0];
    rz_v = nr[1];
    done_u = done_u || nr[2] <= tol2_u;
    done_v = done_v || nr[3] <= tol2_v;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      pu[e] = ru[e] * inv[e] + beta_u * pu[e];
      pv[e] = rv[e] * inv[e] + beta_v * pv[e];
    }
  }
  wave_sync_lds();  // every gather of x0 / p done before x is written back
#pragma unroll
  for (int e = 0; e < E; ++e) {
    if (act[e]) {
      L.xu[lane + 64 * e] = xu[e];
      L.xv[lane + 64 * e] = xv[e];
    }
  }
  wave_sync_lds();
}

template <typename R>
struct SpringArgs {
  R* u;
  R* v;
  const uint8_t* flags;
  const int32_t* count;
  int m, n, sz;
  ReplaceWork w;
};
template <typename R>
struct SpringBody {
  using Args = SpringArgs<R>;
  static __device__ void run(const Args& a, VB vb);
};
template <typename R>
__global__ __launch_bounds__(kSpringLdsThreads) void spring_lds_kernel(SpringArgs<R> a) {
  SpringBody<R>::run(a, vb_native());
}
template <typename R>
__device__ void SpringBody<R>::run(const Args& sa, VB) {
  R* u = sa.u;
  R* v = sa.v;
  const uint8_t* flags = sa.flags;
  const int32_t* count = sa.count;
  const int m = sa.m, n = sa.n, sz = sa.sz;
  const ReplaceWork& w = sa.w;
  constexpr int T = kSpringLdsThreads;
  const int N = m * n;
  const int nv = *count;
  if (nv <= 0 || nv >= N) return;
  __shared__ SpringLds L;
  const int tid = threadIdx.x;

  // row-major compaction: thread tid owns a contiguous run of 16-flag chunks
  const int nchunk = (N + 15) / 16;
  const int per = (nchunk + T - 1) / T;
  const int c0 = tid * per, c1 = min(nchunk, c0 + per);
  const uint4* f4 = reinterpret_cast<const uint4*>(flags);
  auto chunk_words = [&](int c, uint32_t (&words)[4]) {
    if (16 * c + 16 <= N) {
      const uint4 q = f4[c];
      words[0] = q.x;
      words[1] = q.y;
      words[2] = q.z;
      words[3] = q.w;
    } else {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        uint32_t wv = 0;
#pragma unroll
        for (int bb = 0; bb < 4; ++bb) {
          const int k = 16 * c + 4 * q + bb;
          if (k < N) wv |= (uint32_t)flags[k] << (8 * bb);
        }
        words[q] = wv;
      }
    }
  };
  // up to kHeld chunks per thread are loaded at once and kept for the
  // second pass (fields up to 65536 vectors); larger fields load in a loop
  constexpr int kHeld = 16;
  const bool held = per <= kHeld;
  uint32_t hw[kHeld][4];
  int local = 0;
  if (held) {
#pragma unroll
    for (int i = 0; i < kHeld; ++i) {
      if (c0 + i < c1) chunk_words(c0 + i, hw[i]);
      else hw[i][0] = hw[i][1] = hw[i][2] = hw[i][3] = 0;
    }
#pragma unroll
    for (int i = 0; i < kHeld; ++i) local += __popc(hw[i][0]) + __popc(hw[i][1]) + __popc(hw[i][2]) + __popc(hw[i][3]);
  } else {
    for (int c = c0; c < c1; ++c) {
      uint32_t words[4];
      chunk_words(c, words);
      local += __popc(words[0]) + __popc(words[1]) + __popc(words[2]) + __popc(words[3]);  // 0/1 bytes
    }
  }
  // exclusive prefix of the counts: wave scans, then the wave totals
  const int lane = tid & 63, wave = tid >> 6;
  int incl = local;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int y = __shfl_up(incl, off, 64);
    if (lane >= off) incl += y;
  }
  if (lane == 63) L.scan[wave] = incl;
  __syncthreads();
  int pos = incl - local, total = 0;
#pragma unroll
  for (int k = 0; k < T / 64; ++k) {
    pos += k < wave ? L.scan[k] : 0;
    total += L.scan[k];
  }
  const bool in_lds = total <= kSpringLds;
  if (in_lds && held) {
    // visit the set flag bytes only (zero words past the end)
#pragma unroll
    for (int i = 0; i < kHeld; ++i) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        uint32_t wv = hw[i][q];
        while (wv) {
          const int byte = (__builtin_ctz(wv) >> 3);
          L.list[pos++] = 16 * (c0 + i) + 4 * q + byte;
          wv &= ~(0xffu << (8 * byte));
        }
      }
    }
  }
  for (int c = (in_lds && held) ? c1 : c0; c < c1; ++c) {
    uint32_t words[4];
    chunk_words(c, words);
#pragma unroll
    for (int b = 0; b < 16; ++b) {
      const int k = 16 * c + b;
      const bool out = k < N && ((words[b >> 2] >> (8 * (b & 3))) & 0xff);
      if (in_lds) {
        if (out) L.list[pos++] = k;
      } else if (k < N) {
        if (out) {
          w.slot[k] = pos;
          w.list[pos] = k;
          ++pos;
        } else {
          w.slot[k] = -1;
        }
      }
    }
  }
  __syncthreads();
  // With links sz > 1 apart the system splits by residue class (row mod sz,
  // column mod sz).  A class whose every vector is an outlier, with no
  // member in rows [m - sz, sz) or columns [n - sz, sz) (where the
  // position-based diagonal exceeds the link count), is a singular block:
  // its vectors are unresolved and keep their values.
  const int32_t* olist = in_lds ? L.list : w.list;
  if (sz > 1) {
    for (int c = tid; c < sz * sz; c += T) L.cls[c] = 0;
    __syncthreads();
    for (int q = tid; q < total; q += T) {
      const int k = olist[q];
      atomicAdd(&L.cls[(k / n) % sz * sz + (k % n) % sz], 1);
    }
    __syncthreads();
    for (int c = tid; c < sz * sz; c += T) {
      const int a = c / sz, b = c % sz;
      const int members = ((m - a + sz - 1) / sz) * ((n - b + sz - 1) / sz);  // sz <= m, n: a < m, b < n
      L.cls[c] = (a < m - sz && b < n - sz && L.cls[c] == members) ? 1 : 0;
    }
    __syncthreads();
  }
  const int32_t* cls = sz > 1 ? L.cls : nullptr;
  if (!in_lds) {
    spring_cg_global(u, v, m, n, sz, total, w, L.red, cls);
    return;
  }
  const int nvl = total;

  // neighbour codes (slot, -1 valid, -2 outside) and x0 = b / n_k
  const FlagOutlier is_out{flags};
  if (tid == 0) {
    L.xu[kSpringLds] = 0.0;
    L.xv[kSpringLds] = 0.0;
    L.pu[kSpringLds] = 0.0;
    L.pv[kSpringLds] = 0.0;
  }
  auto slot_of = [&](int l) {  // l is in the (ascending) list
    int lo = 0, hi = nvl - 1;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (L.list[mid] < l) lo = mid + 1;
      else hi = mid;
    }
    return lo;
  };
  for (int s = tid; s < nvl; s += T) {
    const int k = L.list[s];
    const int i = k / n, j = k % n;
    const int ni[4] = {i, i, i + sz, i - sz};
    const int nj[4] = {j - sz, j + sz, j, j};
    double bu = 0.0, bv = 0.0;
    const int pk = spring_diag(i, j, m, n, sz);
    bool nout[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {  // flags of the (clamped) neighbours, loaded together
      const int l = min(max(ni[e], 0), m - 1) * n + min(max(nj[e], 0), n - 1);
      nout[e] = is_out(l);
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      int16_t code = -2;
      if (ni[e] >= 0 && ni[e] < m && nj[e] >= 0 && nj[e] < n)
        code = nout[e] ? (int16_t)slot_of(ni[e] * n + nj[e]) : (int16_t)-1;
      L.nb[s][e] = code;
    }
    bool iu, iv;
    spring_rhs(u, v, is_out, m, n, pk, ni, nj, &bu, &bv, &iu, &iv);
    const bool sing = cls && cls[i % sz * sz + j % sz];
    L.iso[s] = (uint8_t)((iu || sing ? 1 : 0) | (iv || sing ? 2 : 0));
    L.nk[s] = (uint8_t)pk;
    L.xu[s] = bu / pk;
    L.xv[s] = bv / pk;
  }
  __syncthreads();
  if (nvl <= 256) {
    if (tid >= 64) return;  // wave 0 finishes alone
    bool cg = true;
    if (!cls) {
      cg = spring_clusters(L, nvl);
    } else {  // singular residue classes possible: CG alone (b = n_k x)
      for (int s = tid; s < nvl; s += 64) {
        L.ru[s] = L.xu[s] * (double)L.nk[s];
        L.rv[s] = L.xv[s] * (double)L.nk[s];
      }
      wave_sync_lds();
    }
    if (cg) spring_cg_wave(L, nvl);
  } else {
    spring_cg<T>(L, nvl);
  }
  // vectors with no outlier and no finite neighbour keep their value
  // (nan_isolated(), found while the right-hand sides were formed)
  for (int s = tid; s < nvl; s += (nvl <= 256 ? 64 : T)) {
    const int k = L.list[s];
    const int iso = L.iso[s];
    if (!(iso & 1)) u[k] = (R)L.xu[s];
    if (!(iso & 2)) v[k] = (R)L.xv[s];
  }
}

struct FillArgs {
  int32_t* p;
  int32_t value;
  int count;
};
struct FillBody {
  using Args = FillArgs;
  static __device__ void run(const Args& a, VB b) {
    const int k = b.bx * blockDim.x + threadIdx.x;
    if (k < a.count) a.p[k] = a.value;
  }
};
__global__ __launch_bounds__(256) void fill_i32_kernel(FillArgs a) { FillBody::run(a, vb_native()); }

template <typename R>
struct ScaleArgs {
  R* u;
  R* v;
  int count;
  R factor;
};
template <typename R>
struct ScaleBody {
  using Args = ScaleArgs<R>;
  static __device__ void run(const Args& a, VB b) {
    const int k = b.bx * blockDim.x + threadIdx.x;
    if (k < a.count) {
      a.u[k] *= a.factor;
      a.v[k] *= a.factor;
    }
  }
};
template <typename R>
__global__ __launch_bounds__(256) void scale_kernel(ScaleArgs<R> a) { ScaleBody<R>::run(a, vb_native()); }
__global__ void scale_div_kernel(double* u, double* v, int count, double mul, double div) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < count) {
    u[k] = u[k] * mul / div;
    v[k] = v[k] * mul / div;
  }
}
template <typename R>
struct CopyArgs {
  R* dst;
  const R* src;
  int count;
};
template <typename R>
struct CopyBody {
  using Args = CopyArgs<R>;
  static __device__ void run(const Args& a, VB b) {
    const int k = b.bx * blockDim.x + threadIdx.x;
    if (k < a.count) a.dst[k] = a.src[k];
  }
};
template <typename R>
__global__ __launch_bounds__(256) void copy_kernel(CopyArgs<R> a) { CopyBody<R>::run(a, vb_native()); }
__global__ void copy_u8_kernel(uint8_t* dst, const uint8_t* src, int count) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < count) dst[k] = src[k];
}

// PIVGPU.mask_frames (gpu_process.py:682-687): masked pixels of both frames
// set to 0, in place, 16 pixels per thread.
template <typename T>
__global__ __launch_bounds__(256) void mask_frames_kernel(T* a, T* b, const uint8_t* mask, size_t npx) {
  const size_t p0 = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) * 16;
  if (p0 >= npx) return;
  if (p0 + 16 <= npx) {
    const uint4 q = *reinterpret_cast<const uint4*>(mask + p0);
    if ((q.x | q.y | q.z | q.w) == 0) return;
    const uint32_t words[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      if ((words[e >> 2] >> (8 * (e & 3))) & 0xff) {
        a[p0 + e] = T(0);
        b[p0 + e] = T(0);
      }
    }
  } else {
    for (size_t p = p0; p < npx; ++p)
      if (mask[p]) {
        a[p] = T(0);
        b[p] = T(0);
      }
  }
}

// Masked windows (CorrelationGPU with field_mask, gpu_process.py:1296-1310,
// 1338-1343): displacement NaN, signal-to-noise 0.
template <typename R>
__global__ __launch_bounds__(256) void apply_field_mask_kernel(R* u, R* v, R* s2n, const uint8_t* fmask, int count) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < count && fmask[k]) {
    u[k] = vnan<R>();
    v[k] = vnan<R>();
    if (s2n) s2n[k] = R(0);
  }
}

thread_local char g_detail[256] = "";


// ---------------------------------------------------------------------------
// Full validation and median/mean replacement (ValidationGPU.__call__,
// gpu_validation.py:84-164; ReplacementGPU, :294-415).
// ---------------------------------------------------------------------------

// Gathers the up-to-8 in-bounds, non-NaN 3x3 neighbours of (i, j) of f
// (optionally leaving out flagged ones) into s[0..cnt), +inf after.
template <typename R>
__device__ __forceinline__ int gather8(const R* f, int m, int n, int i, int j, const uint8_t* skip, R (&s)[8]) {
  int c = 0;
#pragma unroll
  for (int q = 0; q < 8; ++q) s[q] = (R)INFINITY;
#pragma unroll
  for (int di = -1; di <= 1; ++di) {
#pragma unroll
    for (int dj = -1; dj <= 1; ++dj) {
      if (di == 0 && dj == 0) continue;
      const int ii = i + di, jj = j + dj;
      if (ii < 0 || ii >= m || jj < 0 || jj >= n) continue;
      const int idx = ii * n + jj;
      if (skip && skip[idx]) continue;
      const R a = f[idx];
      if (isnan(a)) continue;
#pragma unroll
      for (int q = 0; q < 8; ++q) s[q] = (q == c) ? a : s[q];
      ++c;
    }
  }
  return c;
}

// Median, MAD and mean of the gathered neighbours (NaN when there are none).
template <typename R>
__device__ __forceinline__ void stats8(R (&s)[8], int cnt, R* med, R* mad, R* mean) {
  R sum = R(0);
#pragma unroll
  for (int q = 0; q < 8; ++q) sum += (q < cnt) ? s[q] : R(0);
  *mean = cnt > 0 ? sum / (R)cnt : vnan<R>();
  const R md = median8(s, cnt);
  *med = md;
  if (mad) {
    R d[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) d[q] = (q < cnt) ? vabs(s[q] - md) : (R)INFINITY;
    *mad = median8(d, cnt);
  }
}

// Neighbourhoods wider than 3x3 (half-width 2..kMaxNeighbourhoodSize).  A
// neighbour is addressed by its index e in the (2 size + 1)^2 block, row
// major; the centre, positions outside the field, skipped (flagged)
// positions and NaN values read as NaN and take no part, like the
// reference's NaN padding and NaN centre (gpu_validation.py:194-205,
// 393-398).  Order statistics come from rank counting -- the value whose
// rank range [#less, #less + #equal) holds the wanted rank -- so no
// per-thread array is needed.
template <typename R>
struct Nbhd {
  const R* f;
  const uint8_t* skip;
  int m, n, i, j, size, K;
  __device__ R at(int e) const {
    const int w = 2 * size + 1;
    const int di = e / w - size, dj = e % w - size;
    const int ii = i + di, jj = j + dj;
    if ((di == 0 && dj == 0) || ii < 0 || ii >= m || jj < 0 || jj >= n) return vnan<R>();
    const int idx = ii * n + jj;
    if (skip && skip[idx]) return vnan<R>();
    return f[idx];
  }
};

// Values of ranks r0 <= r1 among the non-NaN values of get(0..K-1); DEV:
// rank |x - centre| instead of x.
template <bool DEV, typename R, typename G>
__device__ void two_ranks(const G& g, R centre, int r0, int r1, R* y0, R* y1) {
  R a0 = R(0), a1 = R(0);
  for (int e = 0; e < g.K; ++e) {
    R x = g.at(e);
    if (isnan(x)) continue;
    if (DEV) x = vabs(x - centre);
    int less = 0, eq = 0;
    for (int q = 0; q < g.K; ++q) {
      R y = g.at(q);
      if (DEV) y = vabs(y - centre);  // NaN stays NaN and compares false
      less += (y < x) ? 1 : 0;
      eq += (y == x) ? 1 : 0;
    }
    if (less <= r0 && r0 < less + eq) a0 = x;
    if (less <= r1 && r1 < less + eq) a1 = x;
  }
  *y0 = a0;
  *y1 = a1;
}

// nanmedian (the mean of the two middle values for an even count, NaN for
// none), MAD (median of |x - median|) and nanmean of the neighbourhood.
template <typename R, typename G>
__device__ void wide_stats(const G& g, R* med, R* mad, R* mean) {
  R sum = R(0);
  int cnt = 0;
  for (int e = 0; e < g.K; ++e) {
    const R x = g.at(e);
    if (!isnan(x)) {
      sum += x;
      ++cnt;
    }
  }
  if (cnt == 0) {
    *med = *mean = vnan<R>();
    if (mad) *mad = vnan<R>();
    return;
  }
  *mean = sum / (R)cnt;
  R y0, y1;
  two_ranks<false>(g, R(0), (cnt - 1) / 2, cnt / 2, &y0, &y1);
  const R md = (cnt & 1) ? y0 : R(0.5) * (y0 + y1);
  *med = md;
  if (mad) {
    two_ranks<true>(g, md, (cnt - 1) / 2, cnt / 2, &y0, &y1);
    *mad = (cnt & 1) ? y0 : R(0.5) * (y0 + y1);
  }
}

template <typename R, bool WIDE>
struct ValidateBody {
  using Args = ValidateArgsT<R>;
  static __device__ void run(const Args& a, VB vb);
};
template <typename R, bool WIDE>
__global__ __launch_bounds__(256) void validate_kernel(ValidateArgsT<R> a) {
  ValidateBody<R, WIDE>::run(a, vb_native());
}
template <typename R, bool WIDE>
__device__ void ValidateBody<R, WIDE>::run(const Args& a, VB vb) {
  const int k = vb.bx * blockDim.x + threadIdx.x;
  bool flag = false;
  if (k < a.m * a.n) {
    if ((!a.only || a.only[k]) && !(a.fmask && a.fmask[k])) {
      const int i = k / a.n, j = k % a.n;
      const R uk = a.u[k], vk = a.v[k];
      if (a.s2n && !isnan(a.s2n_tol)) flag = flag || (a.s2n[k] < a.s2n_tol);
      const bool nb = !isnan(a.median_tol) || !isnan(a.mad_tol) || !isnan(a.mean_tol);
      if (nb) {
        R medu, madu, meanu, medv, madv, meanv;
        if constexpr (WIDE) {
          const int K = (2 * a.size + 1) * (2 * a.size + 1);
          wide_stats(Nbhd<R>{a.u, nullptr, a.m, a.n, i, j, a.size, K}, &medu, isnan(a.mad_tol) ? nullptr : &madu,
                     &meanu);
          wide_stats(Nbhd<R>{a.v, nullptr, a.m, a.n, i, j, a.size, K}, &medv, isnan(a.mad_tol) ? nullptr : &madv,
                     &meanv);
        } else {
          R s[8];
          int c = gather8(a.u, a.m, a.n, i, j, nullptr, s);
          stats8(s, c, &medu, isnan(a.mad_tol) ? nullptr : &madu, &meanu);
          c = gather8(a.v, a.m, a.n, i, j, nullptr, s);
          stats8(s, c, &medv, isnan(a.mad_tol) ? nullptr : &madv, &meanv);
        }
        if (!isnan(a.median_tol))
          flag = flag || (vabs(uk - medu) > a.median_tol) || (vabs(vk - medv) > a.median_tol);
        if (!isnan(a.mad_tol))
          flag = flag || (vabs(uk - medu) > a.mad_tol * madu) || (vabs(vk - medv) > a.mad_tol * madv);
        if (!isnan(a.mean_tol))
          flag = flag || (vabs(uk - meanu) > a.mean_tol) || (vabs(vk - meanv) > a.mean_tol);
      }
      if (!isnan(a.rms_tol) && a.rms) {
        flag = flag || (fabs((double)uk - a.rms[0]) > (double)a.rms_tol * a.rms[1]) ||
               (fabs((double)vk - a.rms[2]) > (double)a.rms_tol * a.rms[3]);
      }
    }
    if (a.extra && a.extra[k]) flag = true;
    a.flags[k] = flag ? 1 : 0;
  }
  const unsigned long long b = __ballot(flag);
  if ((threadIdx.x & 63) == 0 && b) atomicAdd(a.count, (int32_t)__popcll(b));
}

// One 1024-thread block: NaN-mean and population std of u and v in FP64,
// two passes (mean, then sum of squared deviations).
template <typename R>
struct FieldStatsArgs {
  const R* u;
  const R* v;
  int count;
  double* stats;
};
template <typename R>
struct FieldStatsBody {
  using Args = FieldStatsArgs<R>;
  static __device__ void run(const Args& a, VB vb);
};
template <typename R>
__global__ __launch_bounds__(1024) void field_stats_kernel(FieldStatsArgs<R> a) {
  FieldStatsBody<R>::run(a, vb_native());
}
template <typename R>
__device__ void FieldStatsBody<R>::run(const Args& fa, VB) {
  const R* u = fa.u;
  const R* v = fa.v;
  const int count = fa.count;
  double* stats = fa.stats;
  __shared__ double red[2][32];
  __shared__ double mean[2];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  double s[2] = {0.0, 0.0}, c[2] = {0.0, 0.0};
  for (int k = tid; k < count; k += 1024) {
    const R a = u[k], b = v[k];
    if (!isnan(a)) { s[0] += a; c[0] += 1.0; }
    if (!isnan(b)) { s[1] += b; c[1] += 1.0; }
  }
  auto block_sum = [&](double x, int slot) -> double {
    x = wave_sum(x);
    if (lane == 0) red[slot][wave] = x;
    __syncthreads();
    double t = 0.0;
    for (int w = 0; w < 16; ++w) t += red[slot][w];
    __syncthreads();
    return t;
  };
  const double su = block_sum(s[0], 0), sv = block_sum(s[1], 1);
  const double cu = block_sum(c[0], 0), cv = block_sum(c[1], 1);
  if (tid == 0) {
    mean[0] = cu > 0 ? su / cu : NAN;
    mean[1] = cv > 0 ? sv / cv : NAN;
  }
  __syncthreads();
  double q[2] = {0.0, 0.0};
  for (int k = tid; k < count; k += 1024) {
    const R a = u[k], b = v[k];
    if (!isnan(a)) q[0] += ((double)a - mean[0]) * ((double)a - mean[0]);
    if (!isnan(b)) q[1] += ((double)b - mean[1]) * ((double)b - mean[1]);
  }
  const double qu = block_sum(q[0], 0), qv = block_sum(q[1], 1);
  if (tid == 0) {
    stats[0] = mean[0];
    stats[1] = cu > 0 ? sqrt(qu / cu) : NAN;
    stats[2] = mean[1];
    stats[3] = cv > 0 ? sqrt(qv / cv) : NAN;
  }
}

template <typename R>
struct NbrArgs {
  const R* u_in;
  const R* v_in;
  R* u;
  R* v;
  const uint8_t* flags;
  int m, n, size, method, exclude;
};
template <typename R, bool WIDE>
struct NbrBody {
  using Args = NbrArgs<R>;
  static __device__ void run(const Args& a, VB vb);
};
template <typename R, bool WIDE>
__global__ __launch_bounds__(256) void neighbour_replace_kernel(NbrArgs<R> a) {
  NbrBody<R, WIDE>::run(a, vb_native());
}
template <typename R, bool WIDE>
__device__ void NbrBody<R, WIDE>::run(const Args& na, VB vb) {
  const R* u_in = na.u_in;
  const R* v_in = na.v_in;
  R* u = na.u;
  R* v = na.v;
  const uint8_t* flags = na.flags;
  const int m = na.m, n = na.n, size = na.size, method = na.method, exclude = na.exclude;
  const int k = vb.bx * blockDim.x + threadIdx.x;
  if (k >= m * n || !flags[k]) return;
  const int i = k / n, j = k % n;
  const uint8_t* skip = exclude ? flags : nullptr;
  R med, mean;
  if constexpr (WIDE) {
    const int K = (2 * size + 1) * (2 * size + 1);
    wide_stats(Nbhd<R>{u_in, skip, m, n, i, j, size, K}, &med, (R*)nullptr, &mean);
    u[k] = method == 1 ? med : mean;
    wide_stats(Nbhd<R>{v_in, skip, m, n, i, j, size, K}, &med, (R*)nullptr, &mean);
    v[k] = method == 1 ? med : mean;
  } else {
    R s[8];
    int c = gather8(u_in, m, n, i, j, skip, s);
    stats8<R>(s, c, &med, nullptr, &mean);
    u[k] = method == 1 ? med : mean;
    c = gather8(v_in, m, n, i, j, skip, s);
    stats8<R>(s, c, &med, nullptr, &mean);
    v[k] = method == 1 ? med : mean;
  }
}

template <typename R>
struct ResetArgs {
  R* u;
  R* v;
  const R* u0;
  const R* v0;
  const uint8_t* flags;
  uint8_t* nan_flags;
  int count;
};
template <typename R>
struct ResetBody {
  using Args = ResetArgs<R>;
  static __device__ void run(const Args& a, VB vb);
};
template <typename R>
__global__ __launch_bounds__(256) void reset_unresolved_kernel(ResetArgs<R> a) {
  ResetBody<R>::run(a, vb_native());
}
template <typename R>
__device__ void ResetBody<R>::run(const Args& ra, VB vb) {
  R* u = ra.u;
  R* v = ra.v;
  const R* u0 = ra.u0;
  const R* v0 = ra.v0;
  const uint8_t* flags = ra.flags;
  uint8_t* nan_flags = ra.nan_flags;
  const int count = ra.count;
  const int k = vb.bx * blockDim.x + threadIdx.x;
  if (k >= count) return;
  const bool bad = flags[k] && (isnan(u[k]) || isnan(v[k]));
  if (bad) {
    u[k] = u0[k];
    v[k] = v0[k];
  }
  if (nan_flags) nan_flags[k] = bad ? 1 : 0;
}

}  // namespace

const char* launch_detail() { return g_detail; }
void set_launch_detail(const char* what, hipError_t e) {
  snprintf(g_detail, sizeof(g_detail), "%s: %s", what, hipGetErrorString(e));
}

template <typename R>
hipError_t launch_deform(const DeformArgsT<R>& a, hipStream_t stream) {
  if (a.m < 2 || a.n < 2) return hipErrorInvalidValue;
  constexpr int ROWS = 4;
  dim3 grid((a.wd + 255) / 256, (a.ht + ROWS - 1) / ROWS);
  if constexpr (sizeof(R) == 4) {
    const dim3 tgrid((a.wd + 255) / 256, (a.ht + kDeformRows - 1) / kDeformRows);
    switch (a.in_dtype) {
      case kU8: hipLaunchKernelGGL(deform_tile_kernel<uint8_t>, tgrid, dim3(256), 0, stream, a); break;
      case kU16: hipLaunchKernelGGL(deform_tile_kernel<uint16_t>, tgrid, dim3(256), 0, stream, a); break;
      case kF32: hipLaunchKernelGGL(deform_tile_kernel<float>, tgrid, dim3(256), 0, stream, a); break;
      case kF64: hipLaunchKernelGGL((deform_rows_kernel<R, double, ROWS>), grid, dim3(256), 0, stream, a); break;
      default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
  }
  switch (a.in_dtype) {
    case kU8: hipLaunchKernelGGL((deform_rows_kernel<R, uint8_t, ROWS>), grid, dim3(256), 0, stream, a); break;
    case kU16: hipLaunchKernelGGL((deform_rows_kernel<R, uint16_t, ROWS>), grid, dim3(256), 0, stream, a); break;
    case kF32: hipLaunchKernelGGL((deform_rows_kernel<R, float, ROWS>), grid, dim3(256), 0, stream, a); break;
    case kF64: hipLaunchKernelGGL((deform_rows_kernel<R, double, ROWS>), grid, dim3(256), 0, stream, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <typename T>
void launch_deform_batch_t(const DeformArgs* fields, const int* group_prefix, int n_fields, dim3 grid, int threads,
                           hipStream_t stream) {
  if (threads == 128)
    hipLaunchKernelGGL((deform_batch_kernel<T, 128>), grid, dim3(128), 0, stream, fields, group_prefix, n_fields);
  else if (threads == 256)
    hipLaunchKernelGGL((deform_batch_kernel<T, 256>), grid, dim3(256), 0, stream, fields, group_prefix, n_fields);
  else
    hipLaunchKernelGGL((deform_batch_kernel<T, 512>), grid, dim3(512), 0, stream, fields, group_prefix, n_fields);
}

hipError_t launch_deform_batch(const DeformArgs* fields, const int* group_prefix, int n_fields, int total_groups,
                               int max_wd, int in_dtype, int threads, hipStream_t stream) {
  if (total_groups <= 0) return hipSuccess;
  if (threads != 128 && threads != 256 && threads != 512) return hipErrorInvalidValue;
  dim3 grid((max_wd + threads - 1) / threads, total_groups);
  switch (in_dtype) {
    case kU8: launch_deform_batch_t<uint8_t>(fields, group_prefix, n_fields, grid, threads, stream); break;
    case kU16: launch_deform_batch_t<uint16_t>(fields, group_prefix, n_fields, grid, threads, stream); break;
    case kF32: launch_deform_batch_t<float>(fields, group_prefix, n_fields, grid, threads, stream); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// Recording of a launch (FP32 only: batches are FP32).
template <typename R, typename A>
bool recorded(LaunchKind kind, const A& a, dim3 grid, int block) {
  if constexpr (sizeof(R) != 4) return false;
  else return record_launch(kind, a, grid, block);
}

template <typename R>
hipError_t launch_grid_interp(const GridInterpArgsT<R>& a, hipStream_t stream) {
  if (a.m_in < 2 || a.n_in < 2) return hipErrorInvalidValue;
  const int total = a.m_out * a.n_out;
  const dim3 grid((total + 255) / 256);
  if (recorded<R>(LaunchKind::kGridInterp, a, grid, 256)) return hipSuccess;
  hipLaunchKernelGGL(grid_interp_kernel<R>, grid, dim3(256), 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_median_validate(const float* u, const float* v, int m, int n, float tol, uint8_t* flags,
                                  int32_t* count, hipStream_t stream) {
  const int total = m * n;
  hipLaunchKernelGGL(median_kernel, dim3((total + 255) / 256), dim3(256), 0, stream, u, v, m, n, tol, flags,
                     count);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_validate(const ValidateArgsT<R>& a, hipStream_t stream) {
  const int total = a.m * a.n;
  if (a.size < 1 || a.size > kMaxNeighbourhoodSize) return hipErrorInvalidValue;
  const dim3 grid((total + 255) / 256);
  if (recorded<R>(a.size == 1 ? LaunchKind::kValidate : LaunchKind::kValidateWide, a, grid, 256)) return hipSuccess;
  if (a.size == 1)
    hipLaunchKernelGGL((validate_kernel<R, false>), grid, dim3(256), 0, stream, a);
  else
    hipLaunchKernelGGL((validate_kernel<R, true>), grid, dim3(256), 0, stream, a);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_field_stats(const R* u, const R* v, int count, double* stats, hipStream_t stream) {
  const FieldStatsArgs<R> a{u, v, count, stats};
  if (recorded<R>(LaunchKind::kFieldStats, a, dim3(1), 1024)) return hipSuccess;
  hipLaunchKernelGGL(field_stats_kernel<R>, dim3(1), dim3(1024), 0, stream, a);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_neighbour_replace(const R* u_in, const R* v_in, R* u, R* v, const uint8_t* flags, int m, int n,
                                    int size, int method, int exclude, hipStream_t stream) {
  const int total = m * n;
  if (size < 1 || size > kMaxNeighbourhoodSize) return hipErrorInvalidValue;
  const NbrArgs<R> a{u_in, v_in, u, v, flags, m, n, size, method, exclude};
  const dim3 grid((total + 255) / 256);
  if (recorded<R>(size == 1 ? LaunchKind::kNbrReplace : LaunchKind::kNbrReplaceWide, a, grid, 256)) return hipSuccess;
  if (size == 1)
    hipLaunchKernelGGL((neighbour_replace_kernel<R, false>), grid, dim3(256), 0, stream, a);
  else
    hipLaunchKernelGGL((neighbour_replace_kernel<R, true>), grid, dim3(256), 0, stream, a);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_reset_unresolved(R* u, R* v, const R* u0, const R* v0, const uint8_t* flags, uint8_t* nan_flags,
                                   int count, hipStream_t stream) {
  const ResetArgs<R> a{u, v, u0, v0, flags, nan_flags, count};
  const dim3 grid((count + 255) / 256);
  if (recorded<R>(LaunchKind::kReset, a, grid, 256)) return hipSuccess;
  hipLaunchKernelGGL(reset_unresolved_kernel<R>, grid, dim3(256), 0, stream, a);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_spring_replace(R* u, R* v, const uint8_t* flags, const int32_t* count, int m, int n, int size,
                                 const ReplaceWork& w, hipStream_t stream) {
  if (size < 1 || size > kMaxNeighbourhoodSize || size > m || size > n) return hipErrorInvalidValue;
  // the kernel reads the outlier count on the device (graph-capturable, no
  // host sync) and returns at once when there is nothing to solve
  const SpringArgs<R> a{u, v, flags, count, m, n, size, w};
  if (recorded<R>(LaunchKind::kSpring, a, dim3(1), kSpringLdsThreads)) return hipSuccess;
  hipLaunchKernelGGL(spring_lds_kernel<R>, dim3(1), dim3(kSpringLdsThreads), 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_fill_i32(int32_t* p, int32_t value, int count, hipStream_t stream) {
  const FillArgs a{p, value, count};
  const dim3 grid((count + 255) / 256);
  if (record_launch(LaunchKind::kFill, a, grid, 256)) return hipSuccess;
  hipLaunchKernelGGL(fill_i32_kernel, grid, dim3(256), 0, stream, a);
  return hipGetLastError();
}
template <typename R>
hipError_t launch_scale(R* u, R* v, int count, R factor, hipStream_t stream) {
  const ScaleArgs<R> a{u, v, count, factor};
  const dim3 grid((count + 255) / 256);
  if (recorded<R>(LaunchKind::kScale, a, grid, 256)) return hipSuccess;
  hipLaunchKernelGGL(scale_kernel<R>, grid, dim3(256), 0, stream, a);
  return hipGetLastError();
}
hipError_t launch_scale_div(double* u, double* v, int count, double mul, double div, hipStream_t stream) {
  hipLaunchKernelGGL(scale_div_kernel, dim3((count + 255) / 256), dim3(256), 0, stream, u, v, count, mul, div);
  return hipGetLastError();
}
template <typename R>
hipError_t launch_copy(R* dst, const R* src, int count, hipStream_t stream) {
  const CopyArgs<R> a{dst, src, count};
  const dim3 grid((count + 255) / 256);
  if (recorded<R>(LaunchKind::kCopy, a, grid, 256)) return hipSuccess;
  hipLaunchKernelGGL(copy_kernel<R>, grid, dim3(256), 0, stream, a);
  return hipGetLastError();
}

thread_local std::vector<RecordedLaunch>* g_launch_record = nullptr;

hipError_t launch_field_batch(LaunchKind kind, const void* args, const int4* grids, int n, int total_blocks,
                              hipStream_t s) {
  switch (kind) {
    case LaunchKind::kFill: return launch_batched<FillBody, 256>(args, grids, n, total_blocks, s);
    case LaunchKind::kCopy: return launch_batched<CopyBody<float>, 256>(args, grids, n, total_blocks, s);
    case LaunchKind::kScale: return launch_batched<ScaleBody<float>, 256>(args, grids, n, total_blocks, s);
    case LaunchKind::kGridInterp: return launch_batched<GridInterpBody<float>, 256>(args, grids, n, total_blocks, s);
    case LaunchKind::kValidate:
      return launch_batched<ValidateBody<float, false>, 256>(args, grids, n, total_blocks, s);
    case LaunchKind::kValidateWide:
      return launch_batched<ValidateBody<float, true>, 256>(args, grids, n, total_blocks, s);
    case LaunchKind::kFieldStats: return launch_batched<FieldStatsBody<float>, 1024>(args, grids, n, total_blocks, s);
    case LaunchKind::kSpring:
      return launch_batched<SpringBody<float>, kSpringLdsThreads>(args, grids, n, total_blocks, s);
    case LaunchKind::kNbrReplace: return launch_batched<NbrBody<float, false>, 256>(args, grids, n, total_blocks, s);
    case LaunchKind::kNbrReplaceWide:
      return launch_batched<NbrBody<float, true>, 256>(args, grids, n, total_blocks, s);
    case LaunchKind::kReset: return launch_batched<ResetBody<float>, 256>(args, grids, n, total_blocks, s);
    default: return hipErrorInvalidValue;
  }
}
// Up to kCopyMax device-to-device copies in one launch (blockIdx.y picks the
// copy): 16-byte moves where destination, source and size allow, bytes
// otherwise; an entry without a source zero-fills its destination.
__global__ void copy_multi_kernel(CopyList l) {
  const CopySpec c = l.c[blockIdx.y];
  const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  if (!c.src) {  // zero fill
    uint8_t* d = static_cast<uint8_t*>(c.dst);
    for (size_t k = tid; k < c.bytes; k += stride) d[k] = 0;
    return;
  }
  const bool wide = ((reinterpret_cast<uintptr_t>(c.dst) | reinterpret_cast<uintptr_t>(c.src) | c.bytes) & 15) == 0;
  if (wide) {
    uint4* d = static_cast<uint4*>(c.dst);
    const uint4* s = static_cast<const uint4*>(c.src);
    const size_t n = c.bytes / 16;
    for (size_t k = tid; k < n; k += stride) d[k] = s[k];
  } else {
    uint8_t* d = static_cast<uint8_t*>(c.dst);
    const uint8_t* s = static_cast<const uint8_t*>(c.src);
    for (size_t k = tid; k < c.bytes; k += stride) d[k] = s[k];
  }
}

hipError_t launch_copy_multi(const CopyList& l, int n, hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  if (n > kCopyMax) return hipErrorInvalidValue;
  size_t most = 0;
  for (int k = 0; k < n; ++k) most = std::max(most, l.c[k].bytes);
  // enough blocks to stream the largest copy at 16 bytes per thread and step
  const size_t want = (most / 16 + 255) / 256;
  const unsigned bx = (unsigned)std::min<size_t>(std::max<size_t>(want, 1), 2048);
  hipLaunchKernelGGL(copy_multi_kernel, dim3(bx, n), dim3(256), 0, stream, l);
  return hipGetLastError();
}

hipError_t launch_copy_u8(uint8_t* dst, const uint8_t* src, int count, hipStream_t stream) {
  hipLaunchKernelGGL(copy_u8_kernel, dim3((count + 255) / 256), dim3(256), 0, stream, dst, src, count);
  return hipGetLastError();
}

hipError_t launch_mask_frames(void* a, void* b, int in_dtype, const uint8_t* mask, size_t npx, hipStream_t stream) {
  if (((uintptr_t)mask & 15) != 0) return hipErrorInvalidValue;
  const dim3 grid((unsigned)((npx + 16 * 256 - 1) / (16 * 256)));
  switch (in_dtype) {
    case kU8:
      hipLaunchKernelGGL(mask_frames_kernel<uint8_t>, grid, dim3(256), 0, stream, (uint8_t*)a, (uint8_t*)b, mask, npx);
      break;
    case kU16:
      hipLaunchKernelGGL(mask_frames_kernel<uint16_t>, grid, dim3(256), 0, stream, (uint16_t*)a, (uint16_t*)b, mask,
                         npx);
      break;
    case kF32:
      hipLaunchKernelGGL(mask_frames_kernel<float>, grid, dim3(256), 0, stream, (float*)a, (float*)b, mask, npx);
      break;
    case kF64:
      hipLaunchKernelGGL(mask_frames_kernel<double>, grid, dim3(256), 0, stream, (double*)a, (double*)b, mask, npx);
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <typename R>
hipError_t launch_apply_field_mask(R* u, R* v, R* s2n, const uint8_t* fmask, int count, hipStream_t stream) {
  hipLaunchKernelGGL(apply_field_mask_kernel<R>, dim3((count + 255) / 256), dim3(256), 0, stream, u, v, s2n, fmask,
                     count);
  return hipGetLastError();
}

#define PIV_FIELD_INSTANTIATE(R)                                                                                  \
  template hipError_t launch_deform<R>(const DeformArgsT<R>&, hipStream_t);                                      \
  template hipError_t launch_grid_interp<R>(const GridInterpArgsT<R>&, hipStream_t);                             \
  template hipError_t launch_validate<R>(const ValidateArgsT<R>&, hipStream_t);                                  \
  template hipError_t launch_field_stats<R>(const R*, const R*, int, double*, hipStream_t);                      \
  template hipError_t launch_neighbour_replace<R>(const R*, const R*, R*, R*, const uint8_t*, int, int, int, int, \
                                                  int, hipStream_t);                                             \
  template hipError_t launch_reset_unresolved<R>(R*, R*, const R*, const R*, const uint8_t*, uint8_t*, int,       \
                                                 hipStream_t);                                                   \
  template hipError_t launch_spring_replace<R>(R*, R*, const uint8_t*, const int32_t*, int, int, int,              \
                                               const ReplaceWork&, hipStream_t);                                 \
  template hipError_t launch_scale<R>(R*, R*, int, R, hipStream_t);                                              \
  template hipError_t launch_copy<R>(R*, const R*, int, hipStream_t);                                            \
  template hipError_t launch_apply_field_mask<R>(R*, R*, R*, const uint8_t*, int, hipStream_t);
PIV_FIELD_INSTANTIATE(float)
PIV_FIELD_INSTANTIATE(double)
#undef PIV_FIELD_INSTANTIATE

}  // namespace piv
