// Window correlation as dense DFT matrix products on the MFMA units
// (precision mode beside the FP32 FFT path).
//
// For each window pair (a, b) of size W, zero padded to F = 2^ceil(log2 2W):
//   T = a E^T                (W x F, complex)   row DFT of the W nonzero columns
//   A = E T                  (F x F, complex)   column DFT
//   P = conj(A) B
//   U = P G^T, c = Re(G U)   (F x F)           inverse DFT
// with E = D_F[:, pad:pad+W], D_F[k][n] = exp(-2 pi i k n / F) and
// G[x][k] = exp(+2 pi i k x / F).  c is then fftshift-ed and reduced exactly
// like the FFT path (gpu_process.py:1218, 1248-1294).
//
// Every real product runs on v_mfma_f32_16x16x16 (f16 or bf16 inputs, FP32
// accumulation).  In the split modes each FP32 operand x is carried as
// x_hi = round(x), x_lo = round(x - x_hi) and a product is formed from three
// MFMAs (hi*hi + hi*lo + lo*hi), which recovers ~2x the input mantissa.
// Windows and spectra are scaled by their max magnitude before each
// half-precision stage (peak location and the 3-point fits are scale
// invariant), keeping FP16 inside its range.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "piv_kernels.h"

namespace piv {

namespace {

typedef _Float16 half4_t __attribute__((ext_vector_type(4)));
typedef short bf16x4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ short f32_to_bf16_bits(float x) {
  // round to nearest even
  uint32_t u = __builtin_bit_cast(uint32_t, x);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}
__device__ __forceinline__ float bf16_bits_to_f32(short b) {
  return __builtin_bit_cast(float, ((uint32_t)(uint16_t)b) << 16);
}

// One operand of a real batched product: element (r, c) of batch b lives at
// p[b*sb + r*sr + c*sc] (FP32).
struct Opnd {
  const float* p;
  long sb, sr, sc;
};

// C[b] = sum_t sign_t * A_t[b] B_t[b]   (M x N, K inner), up to two terms.
struct DftGemm {
  Opnd a[2], b[2];
  float sign[2];
  int nterms;
  float* c;
  long c_sb, c_sr;
  int M, N, K;
  int batches;
};

template <int PREC>  // 0: fp16 split, 1: bf16 split, 2: fp16, 3: bf16
__device__ __forceinline__ float4_t mfma_split(const float (&av)[4], const float (&bv)[4], float4_t acc) {
  if constexpr (PREC == 0 || PREC == 2) {
    half4_t ah, bh, al, bl;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      ah[i] = (_Float16)av[i];
      bh[i] = (_Float16)bv[i];
      al[i] = (_Float16)(av[i] - (float)ah[i]);
      bl[i] = (_Float16)(bv[i] - (float)bh[i]);
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x16f16(ah, bh, acc, 0, 0, 0);
    if constexpr (PREC == 0) {
      acc = __builtin_amdgcn_mfma_f32_16x16x16f16(ah, bl, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_16x16x16f16(al, bh, acc, 0, 0, 0);
    }
  } else {
    bf16x4_t ah, bh, al, bl;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      ah[i] = f32_to_bf16_bits(av[i]);
      bh[i] = f32_to_bf16_bits(bv[i]);
      al[i] = f32_to_bf16_bits(av[i] - bf16_bits_to_f32(ah[i]));
      bl[i] = f32_to_bf16_bits(bv[i] - bf16_bits_to_f32(bh[i]));
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bh, acc, 0, 0, 0);
    if constexpr (PREC == 1) {
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bl, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(al, bh, acc, 0, 0, 0);
    }
  }
  return acc;
}

// One wave per 16 x 16 output tile of one batch.  Operand layout of
// v_mfma_f32_16x16x16: lane l holds A[l%16][4*(l/16)+e] and
// B[4*(l/16)+e][l%16] (e = 0..3); the result D[4*(l/16)+e][l%16].
template <int PREC>
__global__ __launch_bounds__(256) void dft_gemm_kernel(DftGemm g) {
  const int lane = threadIdx.x & 63;
  const long wave = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int tm = g.M / 16, tn = g.N / 16;
  const long tiles = (long)tm * tn;
  if (wave >= tiles * g.batches) return;
  const int b = (int)(wave / tiles);
  const int t = (int)(wave % tiles);
  const int i0 = (t / tn) * 16, j0 = (t % tn) * 16;
  const int r = lane & 15, kq = (lane >> 4) * 4;
  float4_t acc = {0.f, 0.f, 0.f, 0.f};
  for (int term = 0; term < g.nterms; ++term) {
    const Opnd A = g.a[term], B = g.b[term];
    const float sg = g.sign[term];
    const float* ap = A.p + b * A.sb + (long)(i0 + r) * A.sr;
    const float* bp = B.p + b * B.sb + (long)(j0 + r) * B.sc;
    for (int k0 = 0; k0 < g.K; k0 += 16) {
      float av[4], bv[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        av[e] = sg * ap[(long)(k0 + kq + e) * A.sc];
        bv[e] = bp[(long)(k0 + kq + e) * B.sr];
      }
      acc = mfma_split<PREC>(av, bv, acc);
    }
  }
  float* cp = g.c + b * g.c_sb + j0 + r;
#pragma unroll
  for (int e = 0; e < 4; ++e) cp[(long)(i0 + kq + e) * g.c_sr] = acc[e];
}

// Mean removal and max-magnitude scaling of each window (one wave each).
__global__ __launch_bounds__(64) void dft_normalize_kernel(const float* win, float* out, int ws, int n) {
  const int b = blockIdx.x;
  if (b >= n) return;
  const int lane = threadIdx.x;
  const float* w = win + (long)b * ws * ws;
  float s = 0.f;
  for (int k = lane; k < ws * ws; k += 64) s += w[k];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  const float mean = s / (float)(ws * ws);
  float mx = 0.f;
  for (int k = lane; k < ws * ws; k += 64) mx = fmaxf(mx, fabsf(w[k] - mean));
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off, 64));
  const float inv = mx > 0.f ? 1.0f / mx : 1.0f;
  for (int k = lane; k < ws * ws; k += 64) out[(long)b * ws * ws + k] = (w[k] - mean) * inv;
}

// P = conj(A) B per window, scaled by 1/max|P| so the inverse transform's
// half-precision operands stay in range (the peak and its fit are scale
// invariant).
__global__ __launch_bounds__(256) void dft_product_kernel(const float* ar, const float* ai, const float* br,
                                                          const float* bi, float* pr, float* pi, int F2) {
  const int b = blockIdx.x;
  const long o = (long)b * F2;
  __shared__ float red[4];
  float mx = 0.f;
  for (int k = threadIdx.x; k < F2; k += 256) {
    const float xr = ar[o + k], xi = ai[o + k], yr = br[o + k], yi = bi[o + k];
    const float re = xr * yr + xi * yi, im = xr * yi - xi * yr;
    mx = fmaxf(mx, fmaxf(fabsf(re), fabsf(im)));
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
  __syncthreads();
  const float m = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
  const float inv = m > 0.f ? 1.0f / m : 1.0f;
  for (int k = threadIdx.x; k < F2; k += 256) {
    const float xr = ar[o + k], xi = ai[o + k], yr = br[o + k], yi = bi[o + k];
    pr[o + k] = (xr * yr + xi * yi) * inv;
    pi[o + k] = (xr * yi - xi * yr) * inv;
  }
}

// fftshift + first-occurrence argmax + 3-point fit on the (unshifted) planes.
__global__ __launch_bounds__(64) void dft_peak_kernel(const float* planes, int F, int n, int subpixel, float* u,
                                                      float* v, int32_t* ip_out, int32_t* jp_out, float* shifted) {
  const int b = blockIdx.x;
  if (b >= n) return;
  const int lane = threadIdx.x;
  const float* c = planes + (long)b * F * F;
  const int H = F / 2;
  float best = -INFINITY;
  int bi = 0x7fffffff;
  for (int idx = lane; idx < F * F; idx += 64) {
    // idx runs over the shifted plane: (i, j) <- c[(i+H)%F][(j+H)%F]
    const int i = idx / F, j = idx % F;
    const float val = c[((i + H) % F) * F + (j + H) % F];
    if (shifted) shifted[(long)b * F * F + idx] = val;
    if (val > best || (val == best && idx < bi)) {
      best = val;
      bi = idx;
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const float ov = __shfl_xor(best, off, 64);
    const int oi = __shfl_xor(bi, off, 64);
    if (ov > best || (ov == best && oi < bi)) {
      best = ov;
      bi = oi;
    }
  }
  if (lane != 0) return;
  const int ip = bi / F, jp = bi % F;
  auto at = [&](int i, int j) { return c[((i + H) % F) * F + (j + H) % F]; };
  float isp = (float)ip, jsp = (float)jp;
  if (ip > 0 && ip < F - 1 && jp > 0 && jp < F - 1) {
    const float c0 = at(ip, jp), cd = at(ip - 1, jp), cu = at(ip + 1, jp), cl = at(ip, jp - 1), cr = at(ip, jp + 1);
    const float eps = 1e-7f;
    if (subpixel == 0) {
      if (c0 > 0.f && cd > 0.f && cu > 0.f && cl > 0.f && cr > 0.f) {
        const float lc = logf(c0), lcd = logf(cd), lcu = logf(cu), lcl = logf(cl), lcr = logf(cr);
        isp = (float)ip + 0.5f * (lcd - lcu) / (lcd - 2.0f * lc + lcu + eps);
        jsp = (float)jp + 0.5f * (lcl - lcr) / (lcl - 2.0f * lc + lcr + eps);
      }
    } else if (subpixel == 1) {
      isp = (float)ip + 0.5f * (cd - cu) / (cd - 2.0f * c0 + cu + eps);
      jsp = (float)jp + 0.5f * (cl - cr) / (cl - 2.0f * c0 + cr + eps);
    } else {
      isp = (float)ip + (cu - cd) / (cd + c0 + cu + eps);
      jsp = (float)jp + (cr - cl) / (cl + c0 + cr + eps);
    }
  }
  u[b] = jsp - (float)H;
  v[b] = isp - (float)H;
  if (ip_out) {
    ip_out[b] = ip;
    jp_out[b] = jp;
  }
}

template <int PREC>
hipError_t gemm_launch(const DftGemm& g, hipStream_t s) {
  const long waves = (long)(g.M / 16) * (g.N / 16) * g.batches;
  const long blocks = (waves + 3) / 4;
  hipLaunchKernelGGL(dft_gemm_kernel<PREC>, dim3((unsigned)blocks), dim3(256), 0, s, g);
  return hipGetLastError();
}

hipError_t gemm_any(int prec, const DftGemm& g, hipStream_t s) {
  switch (prec) {
    case 0: return gemm_launch<0>(g, s);
    case 1: return gemm_launch<1>(g, s);
    case 2: return gemm_launch<2>(g, s);
    case 3: return gemm_launch<3>(g, s);
  }
  return hipErrorInvalidValue;
}

}  // namespace

size_t dft_workspace_bytes(int n, int ws, int fsize) {
  const size_t F2 = (size_t)fsize * fsize;
  // normalised windows 2*W^2, T 4*W*F, A/B 4*F^2, P 2*F^2, U 2*F^2, c F^2
  return sizeof(float) * (size_t)n * (2 * ws * ws + 4 * (size_t)ws * fsize + 9 * F2);
}

hipError_t launch_dft_correlate(const DftArgs& d, hipStream_t s) {
  const int n = d.n, W = d.ws, F = d.fsize;
  if (W % 16 || F % 16 || W > F || (F - W) % 2 || n <= 0) return hipErrorInvalidValue;
  const long W2 = (long)W * W, WF = (long)W * F, F2 = (long)F * F;
  float* ws_ = d.workspace;
  float* na = ws_;
  float* nb = na + n * W2;
  float* Tar = nb + n * W2;
  float* Tai = Tar + n * WF;
  float* Tbr = Tai + n * WF;
  float* Tbi = Tbr + n * WF;
  float* Aar = Tbi + n * WF;
  float* Aai = Aar + n * F2;
  float* Abr = Aai + n * F2;
  float* Abi = Abr + n * F2;
  float* Pr = Abi + n * F2;
  float* Pi = Pr + n * F2;
  float* Ur = Pi + n * F2;
  float* Ui = Ur + n * F2;
  float* C = Ui + n * F2;
  hipError_t e;
  hipLaunchKernelGGL(dft_normalize_kernel, dim3(n), dim3(64), 0, s, d.win_a, na, W, n);
  hipLaunchKernelGGL(dft_normalize_kernel, dim3(n), dim3(64), 0, s, d.win_b, nb, W, n);
  // E (F x W complex) and G (F x F complex) tables: er/ei [k*W + n], gr/gi [x*F + k]
  const Opnd Er{d.er, 0, W, 1}, Ei{d.ei, 0, W, 1};
  const Opnd ErT{d.er, 0, 1, W}, EiT{d.ei, 0, 1, W};  // E^T (W x F)
  const Opnd Gr{d.gr, 0, F, 1}, Gi{d.gi, 0, F, 1};
  const Opnd GrT{d.gr, 0, 1, F}, GiT{d.gi, 0, 1, F};
  for (int which = 0; which < 2; ++which) {
    float* src = which ? nb : na;
    float* Tr = which ? Tbr : Tar;
    float* Ti = which ? Tbi : Tai;
    float* Xr = which ? Abr : Aar;
    float* Xi = which ? Abi : Aai;
    const Opnd X{src, W2, W, 1};
    // T = X E^T
    DftGemm g{};
    g.M = W; g.N = F; g.K = W; g.batches = n; g.nterms = 1;
    g.a[0] = X; g.b[0] = ErT; g.sign[0] = 1.f; g.c = Tr; g.c_sb = WF; g.c_sr = F;
    if ((e = gemm_any(d.precision, g, s)) != hipSuccess) return e;
    g.b[0] = EiT; g.c = Ti;
    if ((e = gemm_any(d.precision, g, s)) != hipSuccess) return e;
    // A = E T : Ar = Er Tr - Ei Ti ; Ai = Er Ti + Ei Tr
    const Opnd TrO{Tr, WF, F, 1}, TiO{Ti, WF, F, 1};
    DftGemm h{};
    h.M = F; h.N = F; h.K = W; h.batches = n; h.nterms = 2;
    h.a[0] = Er; h.b[0] = TrO; h.sign[0] = 1.f;
    h.a[1] = Ei; h.b[1] = TiO; h.sign[1] = -1.f;
    h.c = Xr; h.c_sb = F2; h.c_sr = F;
    if ((e = gemm_any(d.precision, h, s)) != hipSuccess) return e;
    h.b[0] = TiO; h.b[1] = TrO; h.sign[1] = 1.f; h.c = Xi;
    if ((e = gemm_any(d.precision, h, s)) != hipSuccess) return e;
  }
  hipLaunchKernelGGL(dft_product_kernel, dim3(n), dim3(256), 0, s, Aar, Aai, Abr, Abi, Pr, Pi, (int)F2);
  // U = P G^T : Ur = Pr Gr^T - Pi Gi^T ; Ui = Pr Gi^T + Pi Gr^T   (P scaled by 1/max|P|)
  {
    const Opnd PrO{Pr, F2, F, 1}, PiO{Pi, F2, F, 1};
    DftGemm g{};
    g.M = F; g.N = F; g.K = F; g.batches = n; g.nterms = 2;
    g.a[0] = PrO; g.b[0] = GrT; g.sign[0] = 1.f;
    g.a[1] = PiO; g.b[1] = GiT; g.sign[1] = -1.f;
    g.c = Ur; g.c_sb = F2; g.c_sr = F;
    if ((e = gemm_any(d.precision, g, s)) != hipSuccess) return e;
    g.b[0] = GiT; g.b[1] = GrT; g.sign[1] = 1.f; g.c = Ui;
    if ((e = gemm_any(d.precision, g, s)) != hipSuccess) return e;
  }
  // c = Gr Ur - Gi Ui
  {
    const Opnd UrO{Ur, F2, F, 1}, UiO{Ui, F2, F, 1};
    DftGemm g{};
    g.M = F; g.N = F; g.K = F; g.batches = n; g.nterms = 2;
    g.a[0] = Gr; g.b[0] = UrO; g.sign[0] = 1.f;
    g.a[1] = Gi; g.b[1] = UiO; g.sign[1] = -1.f;
    g.c = C; g.c_sb = F2; g.c_sr = F;
    if ((e = gemm_any(d.precision, g, s)) != hipSuccess) return e;
  }
  hipLaunchKernelGGL(dft_peak_kernel, dim3(n), dim3(64), 0, s, C, F, n, d.subpixel, d.u, d.v, d.i_peak, d.j_peak,
                     d.planes);
  return hipGetLastError();
}

}  // namespace piv
