// FP64 window correlation for dtype_f = "float64" (CorrelationGPU with
// USE_LONG, gpu_process.py:1021-1365, 1378-1449): windows extracted on the
// pass grid, mean-normalised (:1201-1205, forced on with padding, :1065;
// with an interrogation window inside the search window A is zero outside
// it and normalised over it, :1066-1083, 1194-1199), zero-padded to F x F,
// correlated as irfft2(conj(rfft2(a)) * rfft2(b)) (:1237-1246), fftshift-ed
// (:1218), first-occurrence argmax (:1248-1254), 3-point sub-pixel fit
// (:1256-1294), S2N (:1312-1365), u = j_sp - F/2 (+ predictor), v = i_sp -
// F/2 (:1296-1310) -- every value in double.
//
// One 256-thread workgroup per window (grid-stride over windows).  The two
// real windows travel as one complex signal z = a + i b; the 2-D transform
// is F-point radix-2 FFTs over lines (rows, then columns) staged through a
// small LDS line buffer, F/2 threads per line; the forward row pass only
// transforms the ss nonzero rows and the forward column pass reads only the
// ss nonzero entries.  A and B are split from Z with the k <-> -k mirror and
// P = conj(A) B is written with its Hermitian partner, then the inverse
// columns and rows run on P.  The F x F complex plane lives in LDS for
// F <= 64 and in a per-workgroup global scratch slab for F = 128.
#include <math.h>

#include "piv_kernels.h"

namespace piv {

namespace {

constexpr int kT = 256;          // threads per workgroup
constexpr int kMaxBlocks = 1024;  // resident workgroups (scratch slabs for F = 128)

template <typename T>
__device__ __forceinline__ double ld64(const void* base, size_t idx) {
  return (double)(reinterpret_cast<const T*>(base)[idx]);
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

template <int F>
struct G64 {
  static constexpr int H = F / 2;          // butterflies (threads) per line
  static constexpr int LPB = kT / H;       // lines per round
  static constexpr int LOG = F == 16 ? 4 : (F == 32 ? 5 : (F == 64 ? 6 : 7));
  static constexpr bool kLdsPlane = F <= 64;
  static constexpr size_t kLines = (size_t)LPB * F * sizeof(double2);
  static constexpr size_t kPlane = kLdsPlane ? (size_t)F * F * sizeof(double2) : 0;
  static constexpr size_t kLds = kLines + kPlane + (size_t)H * sizeof(double2);
};

template <int F>
__device__ __forceinline__ int brev(int x) {
  return (int)(__brev((unsigned)x) >> (32 - G64<F>::LOG));
}

// In-place radix-2 DIT transform of the LPB lines (input in bit-reversed
// order): sign -1 forward, +1 inverse, unnormalised.  Every butterfly of a
// stage touches its own pair, so one barrier per stage suffices.
template <int F>
__device__ __forceinline__ void fft_lines(double2* lines, const double2* tw, int sign) {
  constexpr int H = G64<F>::H;
  const int l = threadIdx.x / H, b = threadIdx.x % H;
  double2* x = lines + (size_t)l * F;
  for (int h = 1; h < F; h <<= 1) {
    const int pos = b % h;
    const int i0 = (b / h) * 2 * h + pos, i1 = i0 + h;
    double2 w = tw[pos * (F / (2 * h))];  // exp(-2 pi i pos / 2h)
    if (sign > 0) w.y = -w.y;
    const double2 x0 = x[i0];
    const double2 x1 = cmul(x[i1], w);
    x[i0] = make_double2(x0.x + x1.x, x0.y + x1.y);
    x[i1] = make_double2(x0.x - x1.x, x0.y - x1.y);
    __syncthreads();
  }
}

struct BestD {
  double v;
  int i;
};
__device__ __forceinline__ BestD better(BestD a, BestD b) {
  if (b.v > a.v || (b.v == a.v && b.i < a.i)) return b;
  return a;
}

__device__ __forceinline__ BestD block_best(BestD b, BestD* red) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    BestD o{__shfl_xor(b.v, off, 64), __shfl_xor(b.i, off, 64)};
    b = better(b, o);
  }
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = b;
  __syncthreads();
  BestD r = red[0];
  for (int w = 1; w < kT / 64; ++w) r = better(r, red[w]);
  return r;
}

__device__ __forceinline__ double block_sum(double x, double* red) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = x;
  __syncthreads();
  double r = 0.0;
  for (int w = 0; w < kT / 64; ++w) r += red[w];
  return r;
}

template <int F, typename InT>
__global__ __launch_bounds__(kT) void corr64_kernel(CorrArgs64 a) {
  using G = G64<F>;
  constexpr int H = G::H, LPB = G::LPB;
  extern __shared__ double2 dyn[];
  double2* lines = dyn;
  double2* tw = dyn + (size_t)LPB * F;
  double2* plane = G::kLdsPlane ? tw + H : a.scratch + (size_t)blockIdx.x * F * F;
  __shared__ BestD red_b[kT / 64];
  __shared__ double red_d[kT / 64];
  for (int k = threadIdx.x; k < H; k += kT) tw[k] = make_double2(a.twiddle[2 * k], a.twiddle[2 * k + 1]);
  __syncthreads();
  const int l = threadIdx.x / H, t = threadIdx.x % H;
  const int ss = a.ss, pad = (F - ss) / 2, wp = a.win_pad;
  const int n_win = a.m * a.n;
  const double inv_f2 = 1.0 / ((double)F * (double)F);

  for (int win = blockIdx.x; win < n_win; win += gridDim.x) {
    const int wi = win / a.n, wj = win % a.n;
    const size_t y0 = (size_t)wi * a.spacing + a.y_off, x0 = (size_t)wj * a.spacing + a.x_off;
    // window means
    double ma = 0.0, mb = 0.0;
    if (a.normalize) {
      double sa = 0.0, sb = 0.0;
      for (int e = threadIdx.x; e < ss * ss; e += kT) {
        const int r = e / ss, c = e % ss;
        const size_t idx = (y0 + r) * a.wd + x0 + c;
        if (r >= wp && r < ss - wp && c >= wp && c < ss - wp) sa += ld64<InT>(a.frame_a, idx);
        sb += ld64<InT>(a.frame_b, idx);
      }
      sa = block_sum(sa, red_d);
      sb = block_sum(sb, red_d);
      const int wa = ss - 2 * wp;
      ma = sa / ((double)wa * wa);
      mb = sb / ((double)ss * ss);
    }
    // forward rows of z = (a - ma) + i (b - mb): the ss nonzero rows only
    for (int r0 = 0; r0 < ss; r0 += LPB) {
      const int r = r0 + l;
      for (int x = t; x < F; x += H) {
        double2 z = make_double2(0.0, 0.0);
        const int c = x - pad;
        if (r < ss && c >= 0 && c < ss) {
          const size_t idx = (y0 + r) * a.wd + x0 + c;
          const bool in_a = r >= wp && r < ss - wp && c >= wp && c < ss - wp;
          z.x = in_a ? ld64<InT>(a.frame_a, idx) - ma : 0.0;
          z.y = ld64<InT>(a.frame_b, idx) - mb;
        }
        lines[(size_t)l * F + brev<F>(x)] = z;
      }
      __syncthreads();
      fft_lines<F>(lines, tw, -1);
      if (r < ss)
        for (int x = t; x < F; x += H) plane[(size_t)(pad + r) * F + x] = lines[(size_t)l * F + x];
      __syncthreads();
    }
    // forward columns (rows outside [pad, pad + ss) are zero)
    for (int c0 = 0; c0 < F; c0 += LPB) {
      const int c = c0 + l;
      for (int y = t; y < F; y += H)
        lines[(size_t)l * F + brev<F>(y)] =
            (c < F && y >= pad && y < pad + ss) ? plane[(size_t)y * F + c] : make_double2(0.0, 0.0);
      __syncthreads();
      fft_lines<F>(lines, tw, -1);
      if (c < F)
        for (int y = t; y < F; y += H) plane[(size_t)y * F + c] = lines[(size_t)l * F + y];
      __syncthreads();
    }
    // P = conj(A) B, A = (Z(k) + conj Z(-k)) / 2, B = (Z(k) - conj Z(-k)) / 2i;
    // P(-k) = conj P(k)
    for (int k = threadIdx.x; k < F * F; k += kT) {
      const int ky = k / F, kx = k % F;
      const int km = ((F - ky) & (F - 1)) * F + ((F - kx) & (F - 1));
      if (km < k) continue;
      const double2 Z = plane[k], Zm = plane[km];
      const double2 A = make_double2(0.5 * (Z.x + Zm.x), 0.5 * (Z.y - Zm.y));
      const double2 B = make_double2(0.5 * (Z.y + Zm.y), -0.5 * (Z.x - Zm.x));
      const double2 P = make_double2(A.x * B.x + A.y * B.y, A.x * B.y - A.y * B.x);
      plane[k] = P;
      if (km != k) plane[km] = make_double2(P.x, -P.y);
    }
    __syncthreads();
    // inverse columns, then inverse rows; keep the real part / F^2
    for (int c0 = 0; c0 < F; c0 += LPB) {
      const int c = c0 + l;
      for (int y = t; y < F; y += H)
        lines[(size_t)l * F + brev<F>(y)] = c < F ? plane[(size_t)y * F + c] : make_double2(0.0, 0.0);
      __syncthreads();
      fft_lines<F>(lines, tw, +1);
      if (c < F)
        for (int y = t; y < F; y += H) plane[(size_t)y * F + c] = lines[(size_t)l * F + y];
      __syncthreads();
    }
    for (int r0 = 0; r0 < F; r0 += LPB) {
      const int r = r0 + l;
      for (int x = t; x < F; x += H)
        lines[(size_t)l * F + brev<F>(x)] = r < F ? plane[(size_t)r * F + x] : make_double2(0.0, 0.0);
      __syncthreads();
      fft_lines<F>(lines, tw, +1);
      if (r < F)
        for (int x = t; x < F; x += H) plane[(size_t)r * F + x].x = lines[(size_t)l * F + x].x * inv_f2;
      __syncthreads();
    }
    // fftshift-ed value at (i, j): c[(i + F/2) % F][(j + F/2) % F]
    auto S = [&](int i, int j) -> double { return plane[(size_t)((i + H) & (F - 1)) * F + ((j + H) & (F - 1))].x; };
    BestD best{-INFINITY, 0x7fffffff};
    for (int idx = threadIdx.x; idx < F * F; idx += kT) {
      const double val = S(idx / F, idx % F);
      best = better(best, BestD{val, idx});
      if (a.planes) a.planes[(size_t)win * F * F + idx] = val;
    }
    best = block_best(best, red_b);
    const int ip = best.i / F, jp = best.i % F;
    const double c1 = best.v;
    double s2n = 0.0;
    if (a.s2n) {
      double c2;
      if (a.s2n_method == 0) {
        const int w = a.s2n_size;
        BestD b2{-INFINITY, 0x7fffffff};
        for (int idx = threadIdx.x; idx < F * F; idx += kT) {
          const int i = idx / F, j = idx % F;
          int di = i - ip, dj = j - jp;
          di = di < 0 ? -di : di;
          dj = dj < 0 ? -dj : dj;
          di = di < F - di ? di : F - di;
          dj = dj < F - dj ? dj : F - dj;
          if (di <= w && dj <= w) continue;
          b2 = better(b2, BestD{S(i, j), idx});
        }
        b2 = block_best(b2, red_b);
        c2 = b2.i == 0x7fffffff ? -INFINITY : b2.v;
      } else {
        const double lim = 0.5 * c1;
        double sum = 0.0, cnt = 0.0;
        for (int idx = threadIdx.x; idx < F * F; idx += kT) {
          const double val = S(idx / F, idx % F);
          if (val > 0.0 && (a.s2n_method == 2 || val <= lim)) {
            sum += val;
            cnt += 1.0;
          }
        }
        sum = block_sum(sum, red_d);
        cnt = block_sum(cnt, red_d);
        c2 = cnt > 0.0 ? sum / cnt : NAN;
      }
      s2n = c1 > 0.0 ? (c2 <= 0.0 ? INFINITY : c1 / c2) : 0.0;
    }
    if (threadIdx.x == 0) {
      if (a.s2n) a.s2n[win] = s2n;
      double isp = (double)ip, jsp = (double)jp;
      if (ip > 0 && ip < F - 1 && jp > 0 && jp < F - 1) {
        const double c = S(ip, jp), cd = S(ip - 1, jp), cu = S(ip + 1, jp), cl = S(ip, jp - 1),
                     cr = S(ip, jp + 1);
        const double eps = 1e-7;
        if (a.subpixel == kGaussian) {
          if (c > 0.0 && cd > 0.0 && cu > 0.0 && cl > 0.0 && cr > 0.0) {
            const double lc = log(c), lcd = log(cd), lcu = log(cu), lcl = log(cl), lcr = log(cr);
            isp = (double)ip + 0.5 * (lcd - lcu) / (lcd - 2.0 * lc + lcu + eps);
            jsp = (double)jp + 0.5 * (lcl - lcr) / (lcl - 2.0 * lc + lcr + eps);
          }
        } else if (a.subpixel == kParabolic) {
          isp = (double)ip + 0.5 * (cd - cu) / (cd - 2.0 * c + cu + eps);
          jsp = (double)jp + 0.5 * (cl - cr) / (cl - 2.0 * c + cr + eps);
        } else {
          isp = (double)ip + (cu - cd) / (cd + c + cu + eps);
          jsp = (double)jp + (cr - cl) / (cl + c + cr + eps);
        }
      }
      double uu = jsp - (double)H, vv = isp - (double)H;
      if (a.up) {
        uu += a.up[win];
        vv += a.vp[win];
      }
      a.u[win] = uu;
      a.v[win] = vv;
      if (a.i_peak) {
        a.i_peak[win] = ip;
        a.j_peak[win] = jp;
      }
    }
    __syncthreads();  // the plane is reused by the next window
  }
}

template <int F, typename InT>
hipError_t launch64_ft(const CorrArgs64& a, hipStream_t stream) {
  using G = G64<F>;
  const size_t lds = G::kLds;
  static thread_local bool attr = false;
  if (lds > 65536 && !attr) {
    hipError_t e = hipFuncSetAttribute((const void*)corr64_kernel<F, InT>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)lds);
    if (e != hipSuccess) {
      set_launch_detail("hipFuncSetAttribute(MaxDynamicSharedMemorySize)", e);
      return e;
    }
    attr = true;
  }
  const int blocks = correlate64_blocks(F, a.m * a.n);
  hipLaunchKernelGGL((corr64_kernel<F, InT>), dim3(blocks), dim3(kT), lds, stream, a);
  return hipGetLastError();
}

template <int F>
hipError_t launch64_f(const CorrArgs64& a, hipStream_t stream) {
  switch (a.in_dtype) {
    case kU8: return launch64_ft<F, uint8_t>(a, stream);
    case kU16: return launch64_ft<F, uint16_t>(a, stream);
    case kF32: return launch64_ft<F, float>(a, stream);
    case kF64: return launch64_ft<F, double>(a, stream);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace

int correlate64_blocks(int fsize, int n_windows) {
  const int cap = fsize > 64 ? kMaxBlocks : 8 * kMaxBlocks;
  return n_windows < cap ? (n_windows > 0 ? n_windows : 1) : cap;
}

size_t correlate64_scratch_bytes(int fsize) {
  if (fsize > kMaxSmallFft) return correlate_big_scratch_bytes(fsize, 8, 1 << 30);
  return fsize > 64 ? (size_t)kMaxBlocks * fsize * fsize * sizeof(double2) : 0;
}

hipError_t launch_correlate64(const CorrArgs64& a, hipStream_t stream) {
  if (a.m * a.n <= 0) return hipSuccess;
  if (a.ss > a.fsize || (a.fsize - a.ss) % 2 || (!a.twiddle && a.fsize <= kMaxSmallFft)) return hipErrorInvalidValue;
  if (a.fsize > 64 && !a.scratch) return hipErrorInvalidValue;
  if (a.fsize > kMaxSmallFft) return launch_correlate_big64(a, stream);
  switch (a.fsize) {
    case 16: return launch64_f<16>(a, stream);
    case 32: return launch64_f<32>(a, stream);
    case 64: return launch64_f<64>(a, stream);
    case 128: return launch64_f<128>(a, stream);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace piv
