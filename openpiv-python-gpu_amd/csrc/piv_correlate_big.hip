// Window correlation for FFT sizes 256..1024, in FP32 (dtype_f float32) and
// FP64 (float64): the reference pads each window to 2^ceil(log2(ss n_fft))
// points whatever the size (gpu_process.py:1099-1109), so 128 px windows, or
// 48 / 64 px windows with n_fft = 3, correlate at F = 256.  Same semantics as
// the small-F kernels: windows on the pass grid, mean-normalised (forced on
// with padding, :1065, 1201-1205; A zero outside a shrunk interrogation
// window and normalised over it, :1066-1083, 1194-1199), zero-padded,
// irfft2(conj(rfft2(a)) rfft2(b)) (:1237-1246), fftshift (:1218),
// first-occurrence argmax (:1248-1254), 3-point sub-pixel fit (:1256-1294),
// S2N (:1312-1365), u = j_sp - F/2 (+ predictor), v = i_sp - F/2
// (:1296-1310).
//
// An F x F complex plane (512 KB at F = 256 in FP32) cannot sit in LDS, so
// each 256-thread workgroup owns a slab of device memory for it and walks
// the windows grid-stride.  The two real windows travel as z = a + i b.
// Lines are transformed in batches of LPR lines staged in LDS (64 KB): the
// batch reads LPR consecutive columns of each plane row, so column passes
// load whole 32..256-byte row segments instead of single elements; each
// radix-2 stage of the batch's LPR * F/2 butterflies is spread over the
// workgroup with one barrier per stage.  The forward row pass transforms only
// the ss nonzero rows and the forward column pass reads only them.  The
// twiddles exp(-2 pi i k / F) come from double-precision sincospi.
#include <math.h>

#include "piv_kernels.h"

namespace piv {

namespace {

constexpr int kBT = 256;                          // threads per workgroup
constexpr size_t kBigLineBytes = 65536;           // LDS line batch
constexpr size_t kBigScratchBudget = 512ull << 20;  // plane slabs of one launch

template <typename R>
struct alignas(2 * sizeof(R)) cx {
  R x, y;
};

constexpr int ilog2c(int v) { return v <= 1 ? 0 : 1 + ilog2c(v / 2); }

template <typename R, int F>
struct BigG {
  static constexpr int H = F / 2;
  static constexpr int LOG = ilog2c(F);
  static constexpr int LPR = (int)(kBigLineBytes / (F * sizeof(cx<R>)));  // lines per batch
  static constexpr size_t kLds = kBigLineBytes + (size_t)H * sizeof(cx<R>);
  static_assert(LPR >= 4 && (F % LPR) == 0, "line batch");
};

template <typename T, typename R>
__device__ __forceinline__ R ldpx(const void* base, size_t idx) {
  return (R)(reinterpret_cast<const T*>(base)[idx]);
}

template <int LOG>
__device__ __forceinline__ int brev(int x) {
  return (int)(__brev((unsigned)x) >> (32 - LOG));
}

__device__ __forceinline__ float rlog(float x) { return logf(x); }
__device__ __forceinline__ double rlog(double x) { return log(x); }

// In-place radix-2 DIT transforms of the batch's LPR lines (inputs in
// bit-reversed order), sign -1 forward, +1 inverse, unnormalised.  The
// butterflies of one stage touch disjoint pairs: one barrier per stage.
template <typename R, int F>
__device__ __forceinline__ void fft_batch(cx<R>* lines, const cx<R>* tw, bool inverse) {
  using G = BigG<R, F>;
  constexpr int H = G::H;
  for (int h = 1, lh = 0; h < F; h <<= 1, ++lh) {
    for (int q = threadIdx.x; q < G::LPR * H; q += kBT) {
      const int l = q / H, b = q % H;
      const int pos = b & (h - 1);
      const int i0 = ((b >> lh) << (lh + 1)) + pos, i1 = i0 + h;
      cx<R> w = tw[pos << (G::LOG - 1 - lh)];  // exp(-2 pi i pos / 2h)
      if (inverse) w.y = -w.y;
      cx<R>* x = lines + (size_t)l * F;
      const cx<R> x0 = x[i0], y = x[i1];
      const cx<R> x1{y.x * w.x - y.y * w.y, y.x * w.y + y.y * w.x};
      x[i0] = cx<R>{x0.x + x1.x, x0.y + x1.y};
      x[i1] = cx<R>{x0.x - x1.x, x0.y - x1.y};
    }
    __syncthreads();
  }
}

template <typename R>
struct BestR {
  R v;
  int i;
};
template <typename R>
__device__ __forceinline__ BestR<R> better(BestR<R> a, BestR<R> b) {
  if (b.v > a.v || (b.v == a.v && b.i < a.i)) return b;
  return a;
}

template <typename R>
__device__ __forceinline__ BestR<R> block_best(BestR<R> b, BestR<R>* red) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    BestR<R> o{__shfl_xor(b.v, off, 64), __shfl_xor(b.i, off, 64)};
    b = better(b, o);
  }
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = b;
  __syncthreads();
  BestR<R> r = red[0];
  for (int w = 1; w < kBT / 64; ++w) r = better(r, red[w]);
  return r;
}

__device__ __forceinline__ double block_sum(double x, double* red) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = x;
  __syncthreads();
  double r = 0.0;
  for (int w = 0; w < kBT / 64; ++w) r += red[w];
  return r;
}

// A: CorrArgs (R = float) or CorrArgs64 (R = double); the fields used are
// common to both.
template <typename A, typename R, int F, typename InT>
__global__ __launch_bounds__(kBT) void corr_big_kernel(A a, cx<R>* scratch) {
  using G = BigG<R, F>;
  constexpr int H = G::H, LPR = G::LPR, LOG = G::LOG;
  extern __shared__ __align__(16) unsigned char big_dyn[];
  cx<R>* lines = reinterpret_cast<cx<R>*>(big_dyn);
  cx<R>* tw = lines + (size_t)LPR * F;
  cx<R>* plane = scratch + (size_t)blockIdx.x * F * F;
  __shared__ BestR<R> red_b[kBT / 64];
  __shared__ double red_d[kBT / 64];
  for (int k = threadIdx.x; k < H; k += kBT) {
    double s, c;
    sincospi(-2.0 * (double)k / (double)F, &s, &c);
    tw[k] = cx<R>{(R)c, (R)s};
  }
  __syncthreads();
  const int ss = a.ss, pad = (F - ss) / 2, wp = a.win_pad;
  const int n_win = a.m * a.n;
  const R inv_f2 = (R)1 / ((R)F * (R)F);  // a power of two: exact

  for (int win = blockIdx.x; win < n_win; win += gridDim.x) {
    const int wi = win / a.n, wj = win % a.n;
    const size_t y0 = (size_t)wi * a.spacing + a.y_off, x0 = (size_t)wj * a.spacing + a.x_off;
    R ma = 0, mb = 0;
    if (a.normalize) {
      double sa = 0.0, sb = 0.0;
      for (int e = threadIdx.x; e < ss * ss; e += kBT) {
        const int r = e / ss, c = e % ss;
        const size_t idx = (y0 + r) * a.wd + x0 + c;
        if (r >= wp && r < ss - wp && c >= wp && c < ss - wp) sa += (double)ldpx<InT, R>(a.frame_a, idx);
        sb += (double)ldpx<InT, R>(a.frame_b, idx);
      }
      sa = block_sum(sa, red_d);
      sb = block_sum(sb, red_d);
      const int wa = ss - 2 * wp;
      ma = (R)(sa / ((double)wa * wa));
      mb = (R)(sb / ((double)ss * ss));
    }
    // forward rows of z = (a - ma) + i (b - mb): the ss nonzero rows only
    for (int r0 = 0; r0 < ss; r0 += LPR) {
      for (int q = threadIdx.x; q < LPR * F; q += kBT) {
        const int l = q / F, x = q % F, r = r0 + l, c = x - pad;
        cx<R> z{0, 0};
        if (r < ss && c >= 0 && c < ss) {
          const size_t idx = (y0 + r) * a.wd + x0 + c;
          const bool in_a = r >= wp && r < ss - wp && c >= wp && c < ss - wp;
          z.x = in_a ? ldpx<InT, R>(a.frame_a, idx) - ma : (R)0;
          z.y = ldpx<InT, R>(a.frame_b, idx) - mb;
        }
        lines[(size_t)l * F + brev<LOG>(x)] = z;
      }
      __syncthreads();
      fft_batch<R, F>(lines, tw, false);
      for (int q = threadIdx.x; q < LPR * F; q += kBT) {
        const int l = q / F, x = q % F, r = r0 + l;
        if (r < ss) plane[(size_t)(pad + r) * F + x] = lines[(size_t)l * F + x];
      }
      __syncthreads();
    }
    // forward columns, LPR at a time (rows outside [pad, pad + ss) are zero);
    // q walks the batch row-major so consecutive threads read consecutive
    // columns of one plane row
    for (int c0 = 0; c0 < F; c0 += LPR) {
      for (int q = threadIdx.x; q < LPR * F; q += kBT) {
        const int y = q / LPR, l = q % LPR;
        lines[(size_t)l * F + brev<LOG>(y)] =
            (y >= pad && y < pad + ss) ? plane[(size_t)y * F + c0 + l] : cx<R>{0, 0};
      }
      __syncthreads();
      fft_batch<R, F>(lines, tw, false);
      for (int q = threadIdx.x; q < LPR * F; q += kBT) {
        const int y = q / LPR, l = q % LPR;
        plane[(size_t)y * F + c0 + l] = lines[(size_t)l * F + y];
      }
      __syncthreads();
    }
    // P = conj(A) B with A = (Z(k) + conj Z(-k)) / 2, B = (Z(k) - conj Z(-k)) / 2i;
    // P(-k) = conj P(k)
    for (int k = threadIdx.x; k < F * F; k += kBT) {
      const int ky = k / F, kx = k % F;
      const int km = ((F - ky) & (F - 1)) * F + ((F - kx) & (F - 1));
      if (km < k) continue;
      const cx<R> Z = plane[k], Zm = plane[km];
      const R ax = (R)0.5 * (Z.x + Zm.x), ay = (R)0.5 * (Z.y - Zm.y);
      const R bx = (R)0.5 * (Z.y + Zm.y), by = (R)-0.5 * (Z.x - Zm.x);
      const cx<R> P{ax * bx + ay * by, ax * by - ay * bx};
      plane[k] = P;
      if (km != k) plane[km] = cx<R>{P.x, -P.y};
    }
    __syncthreads();
    // inverse columns, then inverse rows keeping the real part / F^2
    for (int c0 = 0; c0 < F; c0 += LPR) {
      for (int q = threadIdx.x; q < LPR * F; q += kBT) {
        const int y = q / LPR, l = q % LPR;
        lines[(size_t)l * F + brev<LOG>(y)] = plane[(size_t)y * F + c0 + l];
      }
      __syncthreads();
      fft_batch<R, F>(lines, tw, true);
      for (int q = threadIdx.x; q < LPR * F; q += kBT) {
        const int y = q / LPR, l = q % LPR;
        plane[(size_t)y * F + c0 + l] = lines[(size_t)l * F + y];
      }
      __syncthreads();
    }
    for (int r0 = 0; r0 < F; r0 += LPR) {
      for (int q = threadIdx.x; q < LPR * F; q += kBT) {
        const int l = q / F, x = q % F;
        lines[(size_t)l * F + brev<LOG>(x)] = plane[(size_t)(r0 + l) * F + x];
      }
      __syncthreads();
      fft_batch<R, F>(lines, tw, true);
      for (int q = threadIdx.x; q < LPR * F; q += kBT) {
        const int l = q / F, x = q % F;
        plane[(size_t)(r0 + l) * F + x].x = lines[(size_t)l * F + x].x * inv_f2;
      }
      __syncthreads();
    }
    // fftshift-ed value at (i, j): c[(i + F/2) % F][(j + F/2) % F]
    auto S = [&](int i, int j) -> R { return plane[(size_t)((i + H) & (F - 1)) * F + ((j + H) & (F - 1))].x; };
    BestR<R> best{(R)-INFINITY, 0x7fffffff};
    for (int idx = threadIdx.x; idx < F * F; idx += kBT) {
      const R val = S(idx / F, idx % F);
      best = better(best, BestR<R>{val, idx});
      if (a.planes) a.planes[(size_t)win * F * F + idx] = val;
    }
    best = block_best(best, red_b);
    const int ip = best.i / F, jp = best.i % F;
    const R c1 = best.v;
    R s2n = 0;
    if (a.s2n) {
      R c2;
      if (a.s2n_method == 0) {
        // peak2peak: the (2 w + 1)^2 block around the peak (wrapping round
        // the plane edges) left out, gpu_process.py:1318-1320, 1345-1354
        const int w = a.s2n_size;
        BestR<R> b2{(R)-INFINITY, 0x7fffffff};
        for (int idx = threadIdx.x; idx < F * F; idx += kBT) {
          const int i = idx / F, j = idx % F;
          int di = i - ip, dj = j - jp;
          di = di < 0 ? -di : di;
          dj = dj < 0 ? -dj : dj;
          di = di < F - di ? di : F - di;
          dj = dj < F - dj ? dj : F - dj;
          if (di <= w && dj <= w) continue;
          b2 = better(b2, BestR<R>{S(i, j), idx});
        }
        b2 = block_best(b2, red_b);
        c2 = b2.i == 0x7fffffff ? (R)-INFINITY : b2.v;
      } else {
        // peak2mean: mean of the values in (0, peak / 2]; peak2energy: of the
        // positive values (:1322-1327)
        const R lim = (R)0.5 * c1;
        double sum = 0.0, cnt = 0.0;
        for (int idx = threadIdx.x; idx < F * F; idx += kBT) {
          const R val = S(idx / F, idx % F);
          if (val > (R)0 && (a.s2n_method == 2 || val <= lim)) {
            sum += (double)val;
            cnt += 1.0;
          }
        }
        sum = block_sum(sum, red_d);
        cnt = block_sum(cnt, red_d);
        c2 = cnt > 0.0 ? (R)(sum / cnt) : (R)NAN;
      }
      s2n = c1 > (R)0 ? (c2 <= (R)0 ? (R)INFINITY : c1 / c2) : (R)0;
    }
    if (threadIdx.x == 0) {
      if (a.s2n) a.s2n[win] = s2n;
      R isp = (R)ip, jsp = (R)jp;
      if (ip > 0 && ip < F - 1 && jp > 0 && jp < F - 1) {
        const R c = S(ip, jp), cd = S(ip - 1, jp), cu = S(ip + 1, jp), cl = S(ip, jp - 1), cr = S(ip, jp + 1);
        const R eps = (R)1e-7;
        if (a.subpixel == kGaussian) {
          if (c > (R)0 && cd > (R)0 && cu > (R)0 && cl > (R)0 && cr > (R)0) {
            const R lc = rlog(c), lcd = rlog(cd), lcu = rlog(cu), lcl = rlog(cl), lcr = rlog(cr);
            isp = (R)ip + (R)0.5 * (lcd - lcu) / (lcd - (R)2 * lc + lcu + eps);
            jsp = (R)jp + (R)0.5 * (lcl - lcr) / (lcl - (R)2 * lc + lcr + eps);
          }
        } else if (a.subpixel == kParabolic) {
          isp = (R)ip + (R)0.5 * (cd - cu) / (cd - (R)2 * c + cu + eps);
          jsp = (R)jp + (R)0.5 * (cl - cr) / (cl - (R)2 * c + cr + eps);
        } else {
          isp = (R)ip + (cu - cd) / (cd + c + cu + eps);
          jsp = (R)jp + (cr - cl) / (cl + c + cr + eps);
        }
      }
      R uu = jsp - (R)H, vv = isp - (R)H;
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
    __syncthreads();  // the slab is reused by the next window
  }
}

template <typename A, typename R, int F, typename InT>
hipError_t launch_big_ft(const A& a, void* scratch, hipStream_t stream) {
  using G = BigG<R, F>;
  static thread_local bool attr = false;
  if (!attr) {
    hipError_t e = hipFuncSetAttribute((const void*)corr_big_kernel<A, R, F, InT>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)G::kLds);
    if (e != hipSuccess) {
      set_launch_detail("hipFuncSetAttribute(MaxDynamicSharedMemorySize)", e);
      return e;
    }
    attr = true;
  }
  const int blocks = correlate_big_blocks(F, (int)sizeof(R), a.m * a.n);
  hipLaunchKernelGGL((corr_big_kernel<A, R, F, InT>), dim3(blocks), dim3(kBT), G::kLds, stream, a,
                     reinterpret_cast<cx<R>*>(scratch));
  return hipGetLastError();
}

template <typename A, typename R, int F>
hipError_t launch_big_f(const A& a, void* scratch, hipStream_t stream) {
  switch (a.in_dtype) {
    case kU8: return launch_big_ft<A, R, F, uint8_t>(a, scratch, stream);
    case kU16: return launch_big_ft<A, R, F, uint16_t>(a, scratch, stream);
    case kF32: return launch_big_ft<A, R, F, float>(a, scratch, stream);
    case kF64:
      if constexpr (sizeof(R) == 8) return launch_big_ft<A, R, F, double>(a, scratch, stream);
      return hipErrorInvalidValue;
    default: return hipErrorInvalidValue;
  }
}

template <typename A, typename R>
hipError_t launch_big(const A& a, void* scratch, hipStream_t stream) {
  if (a.m * a.n <= 0) return hipSuccess;
  if (!scratch || a.ss < 2 || a.ss > a.fsize || (a.fsize - a.ss) % 2 || a.win_pad < 0 || 2 * a.win_pad >= a.ss) {
    set_launch_detail("correlate (F >= 256): bad geometry or no scratch", hipErrorInvalidValue);
    return hipErrorInvalidValue;
  }
  switch (a.fsize) {
    case 256: return launch_big_f<A, R, 256>(a, scratch, stream);
    case 512: return launch_big_f<A, R, 512>(a, scratch, stream);
    case 1024: return launch_big_f<A, R, 1024>(a, scratch, stream);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace

int correlate_big_blocks(int fsize, int real_bytes, int n_windows) {
  const size_t slab = (size_t)fsize * fsize * 2 * real_bytes;
  const int cap = (int)(kBigScratchBudget / slab);
  return n_windows < cap ? (n_windows > 0 ? n_windows : 1) : cap;
}

size_t correlate_big_scratch_bytes(int fsize, int real_bytes, int n_windows) {
  if (fsize <= kMaxSmallFft) return 0;
  return (size_t)correlate_big_blocks(fsize, real_bytes, n_windows) * fsize * fsize * 2 * real_bytes;
}

hipError_t launch_correlate_big(const CorrArgs& a, hipStream_t stream) {
  return launch_big<CorrArgs, float>(a, a.scratch, stream);
}

hipError_t launch_correlate_big64(const CorrArgs64& a, hipStream_t stream) {
  return launch_big<CorrArgs64, double>(a, a.scratch, stream);
}

}  // namespace piv
