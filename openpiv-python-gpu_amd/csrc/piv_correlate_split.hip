// Window correlation for FFT sizes F = 64 and 128 with every 1-D transform
// split by output residue into four independent sub-transforms.
//
// Same semantics as corr_kernel (CorrelationGPU, gpu_process.py:1021-1310:
// window extraction :1420-1449, mean normalisation :1201-1205, zero padding
// :1222-1235, irfft2(conj(rfft2(a)) * rfft2(b)) :1237-1246, fftshift :1218,
// first-occurrence argmax :1248-1254, 3-point sub-pixel fits :1256-1294,
// u = j_sp - F/2, v = i_sp - F/2 :1296-1310), for launches whose search
// window is at most F/2 and which need neither the S2N epilogue nor an
// interrogation window inside the search window.
//
// Mapping.  A window of W <= F/2 pixels sits at index 0 of its F x F
// transforms (a circular shift common to both frames cancels in
// conj(FA) FB), so a forward 1-D input has F/2 nonzero samples and its
// outputs split by residue r = k mod 4 into four independent length-F/4
// transforms:
//     X[4q + r] = FFT_{F/4}( w^{r n} (x[n] + (-i)^r x[n + F/4]) )[q],
//     w = exp(-2 pi i / F),  n < F/4.
// The inverse transforms split the same way (their outputs by residue, the
// inputs combined four at a time).  Residue r is the wave index, so every
// role-dependent constant is wave-uniform: no cross-lane butterflies and no
// per-lane selects.  Each lane runs whole length-F/4 transforms in registers
// (decimation in time with fused-multiply-add butterflies where the input
// order is free, in frequency where it is not); LDS carries the transposes.
//
//   stage  lane n of residue wave r holds                        LDS
//   R      row n of z = (a - mean) + i (b - mean): Z[4q+r][n]     staging -> Z
//   C      column kx = n of A and B (split from Z columns n and   Z -> regs
//          -n), forward -> conj(FA) FB at ky = 4q + r
//   IC     inverse over q: G_r[y'] w^{-r y'} -> H                 regs -> H
//   comb   R[y' + F/4 t] = sum_r i^{r t} H_r[y']  (t = r)          H -> R
//   IR     rows 2n, 2n+1 as one complex inverse, x = 4q + r        R -> plane
//   peak   first-occurrence argmax, sub-pixel fit                 plane
//
// DC and Nyquist columns are packed into column slot 0 by the residue-0 row
// lanes (stored at kx = 0 and kx = F/2 so the generic A/B split yields
// A0 + i AN and B0 + i BN); their product needs the ky <-> -ky mirror, which
// the window's threads compute through a small LDS exchange.
#include <stdlib.h>

#include "piv_buffer.h"
#include "piv_fft.h"
#include "piv_kernels.h"

namespace piv {

namespace {

template <typename T>
__device__ __forceinline__ float load_px(const void* base, size_t idx) {
  return (float)(reinterpret_cast<const T*>(base)[idx]);
}

template <int F>
struct SplitLayout {
  static constexpr int NR = 4;          // residues = waves per block
  static constexpr int L = F / NR;      // transform length per lane
  static constexpr int HW = F / 2;      // lanes per window and wave
  static constexpr int G = 64 / HW;     // windows per block
  static constexpr int SW = HW + 1;     // staging row stride (complex; odd in 8-byte units)
  static constexpr int ZS = HW + 2;     // Z column stride (complex; odd in 16-byte units)
  static constexpr int HS = F + 2;      // H / R slot stride (complex; odd in 16-byte units)
  static constexpr int SP = F + 1;      // plane row stride (float)
  static constexpr int STAGING = 2 * HW * SW;
  static constexpr int ZF = 2 * F * ZS;
  static constexpr int HF = 2 * HW * HS;
  static constexpr int PF = F * SP;
  static constexpr int M1 = STAGING > ZF ? STAGING : ZF;
  static constexpr int M2 = HF > PF ? HF : PF;
  static constexpr int REGION = ((M1 > M2 ? M1 : M2) + 3) & ~3;  // floats per window
  static constexpr int XQ = 2 * 3 * F;                           // slot-0 exchange (floats)
  // two F = 128 windows (4 waves each) fit a CU's LDS; F = 64 blocks hold
  // two windows in 4 waves
  static constexpr int WAVES = 2;
};

struct SPeak {
  float v;
  int i;
};
__device__ __forceinline__ SPeak sbetter(SPeak a, SPeak b) {
  // larger value wins; equal values -> smaller flat index (first occurrence)
  if (b.v > a.v || (b.v == a.v && b.i < a.i)) return b;
  return a;
}

// (-i)^R v and i^R v
template <int R>
__device__ __forceinline__ c32 rot_m(c32 v) {
  if constexpr ((R & 3) == 0) return v;
  else if constexpr ((R & 3) == 1) return c32{v.y, -v.x};
  else if constexpr ((R & 3) == 2) return c32{-v.x, -v.y};
  else return c32{-v.y, v.x};
}
template <int R>
__device__ __forceinline__ c32 rot_p(c32 v) { return rot_m<4 - (R & 3)>(v); }

// Decimation-in-time FFT of length N on bit-reversed input (x[bitrev(n)]
// holds sample n); natural-order output.  Butterflies with a nontrivial
// twiddle use the fused form u' = u + w v (four FMAs), v' = 2u - u'.
template <int N, int DIR, int S>
struct DitStage {
  template <int B, int J>
  __device__ __forceinline__ static void bfly(c32 (&x)[N]) {
    if constexpr (B < N) {
      if constexpr (J < S) {
        constexpr int k = J * (N / (2 * S));  // twiddle exponent over N
        const c32 u = x[B + J], v = x[B + J + S];
        if constexpr (k == 0) {
          x[B + J] = cadd(u, v);
          x[B + J + S] = csub(u, v);
        } else if constexpr (4 * k == N) {
          const c32 wv = twiddle<k, N, DIR>(v);  // +-i v: a swap
          x[B + J] = cadd(u, wv);
          x[B + J + S] = csub(u, wv);
        } else {
          constexpr float c = (float)cos_frac(k, N);
          constexpr float s = (float)(DIR * sin_frac(k, N));
          c32 t;
          t.x = __builtin_fmaf(c, v.x, __builtin_fmaf(-s, v.y, u.x));
          t.y = __builtin_fmaf(c, v.y, __builtin_fmaf(s, v.x, u.y));
          x[B + J] = t;
          x[B + J + S] = c32{__builtin_fmaf(2.0f, u.x, -t.x), __builtin_fmaf(2.0f, u.y, -t.y)};
        }
        bfly<B, J + 1>(x);
      } else {
        bfly<B + 2 * S, 0>(x);
      }
    }
  }
  __device__ __forceinline__ static void run(c32 (&x)[N]) {
    bfly<0, 0>(x);
    if constexpr (2 * S < N) DitStage<N, DIR, 2 * S>::run(x);
  }
};
template <int N, int DIR>
__device__ __forceinline__ void fft_dit(c32 (&x)[N]) {
  if constexpr (N > 1) DitStage<N, DIR, 1>::run(x);
}

// Runs fn(integral_constant<R>) for the wave-uniform residue (a scalar
// branch: the residue comes from readfirstlane).
template <typename Fn>
__device__ __forceinline__ void with_residue(int r, Fn&& fn) {
  switch (r) {
    case 0: fn(std::integral_constant<int, 0>{}); break;
    case 1: fn(std::integral_constant<int, 1>{}); break;
    case 2: fn(std::integral_constant<int, 2>{}); break;
    default: fn(std::integral_constant<int, 3>{}); break;
  }
}

__device__ __forceinline__ float4 lds_f4(const c32* p) { return *reinterpret_cast<const float4*>(p); }
// Keeps the compiler from hoisting every LDS load of a stage ahead of the
// arithmetic (it would hold them all in registers and spill): loads stay
// in groups of about PERIOD iterations.
template <int I, int PERIOD>
__device__ __forceinline__ void load_fence() {
  if constexpr (I % PERIOD == PERIOD - 1) asm volatile("" ::: "memory");
}
__device__ __forceinline__ void lds_st4(c32* p, c32 a, c32 b) {
  *reinterpret_cast<float4*>(p) = float4{a.x, a.y, b.x, b.y};
}

}  // namespace

size_t correlate_split_lds_bytes(int fsize) {
  return fsize == 128 ? (size_t)(SplitLayout<128>::REGION + SplitLayout<128>::XQ) * SplitLayout<128>::G * 4
                      : (size_t)(SplitLayout<64>::REGION + SplitLayout<64>::XQ) * SplitLayout<64>::G * 4;
}

template <int F, typename InT, bool BATCHED>
__global__ __launch_bounds__(256, SplitLayout<F>::WAVES) void corr_split_kernel(CorrArgs a_in, CorrBatch batch,
                                                                              int n_windows_in) {
  using Ly = SplitLayout<F>;
  constexpr int NR = Ly::NR, L = Ly::L, HW = Ly::HW, G = Ly::G;
  constexpr int SW = Ly::SW, ZS = Ly::ZS, HS = Ly::HS, SP = Ly::SP;
  static_assert(NR == 4, "four residues");

  extern __shared__ float4 smem4[];
  float* smem = reinterpret_cast<float*>(smem4);
  __shared__ float red_f[G][NR][2];
  __shared__ SPeak red_p[G][NR];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int role = __builtin_amdgcn_readfirstlane(tid >> 6);  // residue, wave-uniform
  const int g = lane / HW;   // window slot in the block
  const int n = lane % HW;   // row / column slot / row pair of this lane
  const int t = role * HW + n;  // thread index within the window

  int bid = blockIdx.x;
  if (!BATCHED || batch.xcd) {
    const int nb = gridDim.x;
    const int per = nb / 8;
    if (per > 0 && bid < per * 8) bid = (bid & 7) * per + (bid >> 3);
  }
  CorrArgs a;
  int n_windows;
  if constexpr (BATCHED) {
    int lo = 0, hi = batch.n_fields - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (batch.block_prefix[mid] <= bid) lo = mid;
      else hi = mid - 1;
    }
    a = batch.fields[lo];
    bid -= batch.block_prefix[lo];
    n_windows = a.m * a.n;
  } else {
    a = a_in;
    n_windows = n_windows_in;
  }
  const int win = bid * G + g;
  const bool valid = win < n_windows;
  const int W = a.ss;
  // the predictor of the window, loaded now so the epilogue does not wait
  float pred_u = 0.f, pred_v = 0.f;
  if (a.up && t == 0 && valid) {
    pred_u = a.up[win];
    pred_v = a.vp[win];
  }

  float* region = smem + (size_t)g * Ly::REGION;
  float2* st = reinterpret_cast<float2*>(region);  // staging: (a, b) pixel pairs, row stride SW
  c32* Z = reinterpret_cast<c32*>(region);
  c32* Hb = reinterpret_cast<c32*>(region);
  float* plane = region;
  c32* xq = reinterpret_cast<c32*>(smem + (size_t)G * Ly::REGION + (size_t)g * Ly::XQ);

  // ---- staging: the window's W x W pixels of both frames, centred --------
  {
    const int wrow = valid ? (win / a.n) : 0;
    const int wcol = valid ? (win % a.n) : 0;
    const int row0 = wrow * a.spacing + a.y_off;
    const int col0 = wcol * a.spacing + a.x_off;
    constexpr int PER = (HW * HW) / (2 * F);  // pixels per thread (HW x HW staging area)
    // 2-byte and 4-byte frames inside the 32-bit buffer range: horizontally
    // adjacent pixel pairs per thread (one 4-byte or two merged loads per
    // pair and frame, one paired LDS store), else one pixel per thread
    constexpr bool kPairs = sizeof(InT) == 2 || sizeof(InT) == 4;
    constexpr int HP = HW / 2;        // pixel pairs per staging row
    constexpr int RS = 2 * F / HP;    // staging rows per step of the window's threads
    constexpr int P2 = HW / RS;       // pairs per thread
    static_assert(HW % RS == 0 && (2 * F) % HP == 0, "whole staging rows per step");
    const bool pairs = kPairs && buffer_fits((uint64_t)a.ht * a.wd, sizeof(InT));
    float va[PER], vb[PER];
    float sa = 0.f, sb = 0.f;
    if (pairs) {
      const uint32_t npx = (uint32_t)a.ht * (uint32_t)a.wd;
      const rsrc_t ra = make_rsrc(a.frame_a, npx * (uint32_t)sizeof(InT));
      const rsrc_t rb = make_rsrc(a.frame_b, npx * (uint32_t)sizeof(InT));
      // pair e of this thread: row t / HP + e * RS, columns 2 (t % HP) and + 1
      const int c2 = 2 * (t % HP);
      const uint32_t base = (uint32_t)row0 * (uint32_t)a.wd + (uint32_t)(col0 + min(c2, max(W - 2, 0)));
#pragma unroll
      for (int e = 0; e < P2; ++e) {
        const int r = t / HP + e * RS;
        const uint32_t gi = base + (uint32_t)min(r, W - 1) * (uint32_t)a.wd;
        if constexpr (sizeof(InT) == 2) {
          const uint32_t wa = __builtin_amdgcn_raw_buffer_load_b32(ra, (int)(gi * 2u), 0, 0);
          const uint32_t wb = __builtin_amdgcn_raw_buffer_load_b32(rb, (int)(gi * 2u), 0, 0);
          va[2 * e] = (float)(wa & 0xffffu);
          va[2 * e + 1] = (float)(wa >> 16);
          vb[2 * e] = (float)(wb & 0xffffu);
          vb[2 * e + 1] = (float)(wb >> 16);
        } else {
          va[2 * e] = buf_f32(ra, gi * 4u, 0u);
          va[2 * e + 1] = buf_f32(ra, gi * 4u + 4u, 0u);
          vb[2 * e] = buf_f32(rb, gi * 4u, 0u);
          vb[2 * e + 1] = buf_f32(rb, gi * 4u + 4u, 0u);
        }
      }
#pragma unroll
      for (int e = 0; e < P2; ++e) {
        const int r = t / HP + e * RS;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const bool in = valid && r < W && c2 + h < W;
          va[2 * e + h] = in ? va[2 * e + h] : 0.f;
          vb[2 * e + h] = in ? vb[2 * e + h] : 0.f;
          sa += va[2 * e + h];
          sb += vb[2 * e + h];
        }
      }
    } else {
      if (buffer_fits((uint64_t)a.ht * a.wd, sizeof(InT))) {
        const uint32_t npx = (uint32_t)a.ht * (uint32_t)a.wd;
        const rsrc_t ra = make_rsrc(a.frame_a, npx * (uint32_t)sizeof(InT));
        const rsrc_t rb = make_rsrc(a.frame_b, npx * (uint32_t)sizeof(InT));
        // pixel e of this thread: row t / HW + e * (2F / HW), column t % HW
        const uint32_t base = (uint32_t)row0 * (uint32_t)a.wd + (uint32_t)(col0 + min(t % HW, W - 1));
#pragma unroll
        for (int e = 0; e < PER; ++e) {
          const int r = t / HW + e * (2 * F / HW);
          const uint32_t gi = base + (uint32_t)min(r, W - 1) * (uint32_t)a.wd;
          va[e] = buf_px<InT>(ra, gi);
          vb[e] = buf_px<InT>(rb, gi);
        }
      } else {
#pragma unroll
        for (int e = 0; e < PER; ++e) {
          const int r = t / HW + e * (2 * F / HW);
          const size_t gi = (size_t)(row0 + min(r, W - 1)) * a.wd + (col0 + min(t % HW, W - 1));
          va[e] = load_px<InT>(a.frame_a, gi);
          vb[e] = load_px<InT>(a.frame_b, gi);
        }
      }
      static_assert((2 * F) % HW == 0, "whole staging rows per step");
#pragma unroll
      for (int e = 0; e < PER; ++e) {
        const bool in = valid && t / HW + e * (2 * F / HW) < W && t % HW < W;
        va[e] = in ? va[e] : 0.f;
        vb[e] = in ? vb[e] : 0.f;
        sa += va[e];
        sb += vb[e];
      }
    }
#pragma unroll
    for (int off = HW / 2; off > 0; off >>= 1) {
      sa += __shfl_xor(sa, off, 64);
      sb += __shfl_xor(sb, off, 64);
    }
    if (n == 0) {
      red_f[g][role][0] = sa;
      red_f[g][role][1] = sb;
    }
    __syncthreads();
    sa = red_f[g][0][0] + red_f[g][1][0] + red_f[g][2][0] + red_f[g][3][0];
    sb = red_f[g][0][1] + red_f[g][1][1] + red_f[g][2][1] + red_f[g][3][1];
    // means over the window (forced on: the padding is nonzero,
    // gpu_process.py:1065, 1201-1205)
    const float ma = sa / (float)(W * W), mb = sb / (float)(W * W);
    if (pairs) {
      const int c2 = 2 * (t % HP);
#pragma unroll
      for (int e = 0; e < P2; ++e) {
        const int r = t / HP + e * RS;
        const bool in0 = r < W && c2 < W, in1 = r < W && c2 + 1 < W;
        st[r * SW + c2] = in0 ? float2{va[2 * e] - ma, vb[2 * e] - mb} : float2{0.f, 0.f};
        st[r * SW + c2 + 1] = in1 ? float2{va[2 * e + 1] - ma, vb[2 * e + 1] - mb} : float2{0.f, 0.f};
      }
    } else {
#pragma unroll
      for (int e = 0; e < PER; ++e) {
        const int idx = t + e * 2 * F;
        const int r = idx / HW, c = idx % HW;
        const bool in = r < W && c < W;
        st[r * SW + c] = in ? float2{va[e] - ma, vb[e] - mb} : float2{0.f, 0.f};
      }
    }
  }
  __syncthreads();

  // ---- stage R: row n, residue r: Z[4q + r][n] ------------------------------
  {
    c32 y[L];
    with_residue(role, [&](auto RR) {
      constexpr int R = decltype(RR)::value;
      static_for<0, L>([&](auto CC) {
        constexpr int c = decltype(CC)::value;
        const float2 s0 = st[n * SW + c], s1 = st[n * SW + c + L];
        const c32 z0{s0.x, s0.y}, z1{s1.x, s1.y};
        y[bitrev(c, L)] = twiddle<R * c, F, -1>(cadd(z0, rot_m<R>(z1)));
      });
    });
    __syncthreads();  // staging consumed (Z aliases it)
    fft_dit<L, -1>(y);
    if (role == 0) {
      // pack DC (q = 0) and Nyquist (q = L/2) so that the column split of
      // slot 0 gives (A0 + i AN, B0 + i BN)
      const c32 z0 = y[0], zn = y[L / 2];
      y[0] = c32{z0.x - zn.y, zn.x + z0.y};
      y[L / 2] = c32{z0.x + zn.y, z0.y - zn.x};
    }
    static_for<0, L>([&](auto QQ) {
      constexpr int q = decltype(QQ)::value;
      Z[(size_t)(NR * q + role) * ZS + n] = y[q];
    });
  }
  __syncthreads();

  // ---- stage C: columns kx = n of A and B, residue r ----------------------
  c32 fa[L], fb[L];
  {
    const int kx = n;
    const int km = n == 0 ? HW : F - n;
    const c32* zk = Z + (size_t)kx * ZS;
    const c32* zm = Z + (size_t)km * ZS;
    with_residue(role, [&](auto RR) {
      constexpr int R = decltype(RR)::value;
      static_for<0, L / 2>([&](auto CC) {
        constexpr int c = 2 * decltype(CC)::value;
        const float4 k0 = lds_f4(zk + c), k1 = lds_f4(zk + c + L);
        const float4 m0 = lds_f4(zm + c), m1 = lds_f4(zm + c + L);
        // A = Zk + conj(Zm), B = -i (Zk - conj(Zm)) (each twice the true
        // value); the residue input of A is P + conj(Q) and of B
        // -i (P - conj(Q)), with P = w^{Rc} (Zk[c] + (-i)^R Zk[c + L]) and
        // Q = w^{-Rc} (Zm[c] + i^R Zm[c + L])
        auto put = [&](auto EE, c32 k0e, c32 k1e, c32 m0e, c32 m1e) {
          constexpr int e = c + decltype(EE)::value;
          const c32 p = twiddle<R * e, F, -1>(cadd(k0e, rot_m<R>(k1e)));
          const c32 q = twiddle<R * e, F, +1>(cadd(m0e, rot_p<R>(m1e)));
          fa[bitrev(e, L)] = c32{p.x + q.x, p.y - q.y};
          fb[bitrev(e, L)] = c32{p.y + q.y, q.x - p.x};
        };
        put(std::integral_constant<int, 0>{}, c32{k0.x, k0.y}, c32{k1.x, k1.y}, c32{m0.x, m0.y}, c32{m1.x, m1.y});
        put(std::integral_constant<int, 1>{}, c32{k0.z, k0.w}, c32{k1.z, k1.w}, c32{m0.z, m0.w}, c32{m1.z, m1.w});
        load_fence<decltype(CC)::value, 4>();
      });
    });
  }
  fft_dit<L, -1>(fa);
  fft_dit<L, -1>(fb);
  // slot 0 holds FFT(A0 + i AN) and FFT(B0 + i BN): hand them to the
  // window's threads for the mirror unpack.  The exchange stores frequency
  // ky = NR q + r at (ky mod NR) L + ky / NR, so a residue's values are
  // contiguous and lane 0 moves two per 16-byte access.
  auto xslot = [](int ky) { return (ky & (NR - 1)) * L + (ky >> 2); };
  if (n == 0) {
    static_for<0, L / 2>([&](auto QQ) {
      constexpr int q = 2 * decltype(QQ)::value;
      lds_st4(xq + role * L + q, fa[q], fa[q + 1]);
      lds_st4(xq + F + role * L + q, fb[q], fb[q + 1]);
    });
  }
  // product conj(FA) FB (slot 0's is replaced below)
  static_for<0, L>([&](auto QQ) {
    constexpr int q = decltype(QQ)::value;
    fa[q] = cconjmul(fa[q], fb[q]);
  });
  __syncthreads();  // Z consumed (H aliases it); slot-0 spectra in xq
  if (t < F) {
    // f0 = (q + conj(qm)) / 2, fn = (q - conj(qm)) / (2i); P0 + i PN
    const int ky = t, kmir = (F - t) & (F - 1);
    const c32 qa = xq[xslot(ky)], qam = xq[xslot(kmir)], qb = xq[F + xslot(ky)], qbm = xq[F + xslot(kmir)];
    const c32 fa0{0.5f * (qa.x + qam.x), 0.5f * (qa.y - qam.y)};
    const c32 fan{0.5f * (qa.y + qam.y), 0.5f * (qam.x - qa.x)};
    const c32 fb0{0.5f * (qb.x + qbm.x), 0.5f * (qb.y - qbm.y)};
    const c32 fbn{0.5f * (qb.y + qbm.y), 0.5f * (qbm.x - qb.x)};
    const c32 p0 = cconjmul(fa0, fb0), pn = cconjmul(fan, fbn);
    xq[2 * F + xslot(ky)] = c32{p0.x - pn.y, p0.y + pn.x};
  }
  __syncthreads();
  if (n == 0) {
    static_for<0, L / 2>([&](auto QQ) {
      constexpr int q = 2 * decltype(QQ)::value;
      const float4 v = lds_f4(xq + 2 * F + role * L + q);
      fa[q] = c32{v.x, v.y};
      fa[q + 1] = c32{v.z, v.w};
    });
  }

  // ---- stage IC: inverse over q (natural input, bit-reversed output) -----
  fft_reg<L, +1>(fa);
  {
    c32* hs = Hb + (size_t)n * HS + role * L;
    with_residue(role, [&](auto RR) {
      constexpr int R = decltype(RR)::value;
      static_for<0, L / 2>([&](auto YY) {
        constexpr int y = 2 * decltype(YY)::value;
        const c32 h0 = twiddle<R * y, F, +1>(fa[bitrev(y, L)]);
        const c32 h1 = twiddle<R * (y + 1), F, +1>(fa[bitrev(y + 1, L)]);
        lds_st4(hs + y, h0, h1);
      });
    });
  }
  __syncthreads();

  // ---- combine the four residues: R[y' + L t] = sum_r i^{r t} H_r[y'] ---
  // Wave w takes y' in [w L/4, (w + 1) L/4) and forms all four t: it reads
  // and rewrites only its own slots, so the combine is in place without a
  // barrier between its loads and stores.
  {
    c32* hs = Hb + (size_t)n * HS + role * (L / 4);
    static_for<0, L / 8>([&](auto YY) {
      constexpr int y = 2 * decltype(YY)::value;
      const float4 h0 = lds_f4(hs + y), h1 = lds_f4(hs + L + y), h2 = lds_f4(hs + 2 * L + y),
                   h3 = lds_f4(hs + 3 * L + y);
      c32 r0[2], r1[2], r2[2], r3[2];
      {
        const c32 s02 = cadd(c32{h0.x, h0.y}, c32{h2.x, h2.y}), d02 = csub(c32{h0.x, h0.y}, c32{h2.x, h2.y});
        const c32 s13 = cadd(c32{h1.x, h1.y}, c32{h3.x, h3.y}), d13 = csub(c32{h1.x, h1.y}, c32{h3.x, h3.y});
        r0[0] = cadd(s02, s13);
        r2[0] = csub(s02, s13);
        r1[0] = cadd(d02, rot_p<1>(d13));
        r3[0] = csub(d02, rot_p<1>(d13));
      }
      {
        const c32 s02 = cadd(c32{h0.z, h0.w}, c32{h2.z, h2.w}), d02 = csub(c32{h0.z, h0.w}, c32{h2.z, h2.w});
        const c32 s13 = cadd(c32{h1.z, h1.w}, c32{h3.z, h3.w}), d13 = csub(c32{h1.z, h1.w}, c32{h3.z, h3.w});
        r0[1] = cadd(s02, s13);
        r2[1] = csub(s02, s13);
        r1[1] = cadd(d02, rot_p<1>(d13));
        r3[1] = csub(d02, rot_p<1>(d13));
      }
      lds_st4(hs + y, r0[0], r0[1]);
      lds_st4(hs + L + y, r1[0], r1[1]);
      lds_st4(hs + 2 * L + y, r2[0], r2[1]);
      lds_st4(hs + 3 * L + y, r3[0], r3[1]);
    });
  }
  __syncthreads();

  // ---- stage IR: rows 2n, 2n+1 as one complex inverse, outputs x = 4q + r --
  c32 o[L];
  {
    const c32* rb = Hb + 2 * n;  // slot s of rows (2n, 2n+1) at rb + s * HS
    auto ld = [&](int s) { return lds_f4(rb + (size_t)s * HS); };
    // z[k] = R1[k] + i R2[k] with the Hermitian extension of both rows;
    // slot 0 carries (DC, Nyquist) of each row
    auto z_direct = [](float4 v) { return c32{v.x - v.w, v.y + v.z}; };  // 0 < k < F/2
    auto z_mirror = [](float4 v) { return c32{v.x + v.w, v.z - v.y}; };  // F/2 < k: slot F - k
    // output k reads slots k, k + L, 2L - k, L - k; output L - k reads the
    // same four slots, so the two are formed from one set of loads
    auto out = [&](auto RR, auto KK, c32 z0, c32 z1, c32 z2, c32 z3) {
      constexpr int R = decltype(RR)::value;
      constexpr int k = decltype(KK)::value;
      const c32 s02 = cadd(z0, rot_p<2 * R>(z2));
      const c32 s13 = cadd(z1, rot_p<2 * R>(z3));
      o[bitrev(k, L)] = twiddle<R * k, F, +1>(cadd(s02, rot_p<R>(s13)));
    };
    with_residue(role, [&](auto RR) {
      {
        const float4 s0 = ld(0), sl = ld(L);
        // DC and Nyquist of both rows in slot 0; k = L and k = 3L -> slot L
        out(RR, std::integral_constant<int, 0>{}, c32{s0.x, s0.z}, z_direct(sl), c32{s0.y, s0.w}, z_mirror(sl));
      }
      {
        const float4 sa = ld(L / 2), sb = ld(L + L / 2);
        out(RR, std::integral_constant<int, L / 2>{}, z_direct(sa), z_direct(sb), z_mirror(sb), z_mirror(sa));
      }
      static_for<1, L / 2>([&](auto KK) {
        constexpr int k = decltype(KK)::value;
        const float4 sa = ld(k), sb = ld(k + L), sc = ld(2 * L - k), sd = ld(L - k);
        out(RR, KK, z_direct(sa), z_direct(sb), z_mirror(sc), z_mirror(sd));
        out(RR, std::integral_constant<int, L - k>{}, z_direct(sd), z_direct(sc), z_mirror(sb), z_mirror(sa));
        load_fence<k, 4>();
      });
    });
  }
  fft_dit<L, +1>(o);
  __syncthreads();  // R consumed (the plane aliases it)

  // ---- plane (unscaled: the scale 1/(4 F^2) is a power of two) and peak ---
  SPeak best;
  {
    const int i1 = (2 * n) ^ (F / 2);
    const int i2 = i1 + 1;
    float pm[4] = {-INFINITY, -INFINITY, -INFINITY, -INFINITY};  // partial maxima: no single chain
    static_for<0, L>([&](auto QQ) {
      constexpr int q = decltype(QQ)::value;
      const int jj = (NR * q + role) ^ (F / 2);
      plane[i1 * SP + jj] = o[q].x;
      plane[i2 * SP + jj] = o[q].y;
      pm[q & 3] = fmaxf(pm[q & 3], fmaxf(o[q].x, o[q].y));
    });
    const float zmax = fmaxf(fmaxf(pm[0], pm[1]), fmaxf(pm[2], pm[3]));
    // first occurrence: the lane's values carry codes in flat-index order,
    // code = row * L + (q ^ L/2) (row 0 = i1, 1 = i2; column jj(q) ascends
    // with q ^ L/2 for the wave's residue).  Four independent chains (rows x
    // code halves) walk their codes downwards with compile-time codes; the
    // smallest code wins.
    int fc[4] = {2 * L, 2 * L, 2 * L, 2 * L};
    static_for<0, L / 2>([&](auto KK) {
      constexpr int cl = L / 2 - 1 - decltype(KK)::value;  // lower code half
      constexpr int ch = cl + L / 2;                        // upper code half
      fc[0] = o[cl ^ (L / 2)].x == zmax ? cl : fc[0];
      fc[1] = o[ch ^ (L / 2)].x == zmax ? ch : fc[1];
      fc[2] = o[cl ^ (L / 2)].y == zmax ? L + cl : fc[2];
      fc[3] = o[ch ^ (L / 2)].y == zmax ? L + ch : fc[3];
    });
    const int code = min(min(fc[0], fc[1]), min(fc[2], fc[3]));
    int first = 0x7fffffff;
    if (code < 2 * L) {
      const int q = (code & (L - 1)) ^ (L / 2);
      first = (code < L ? i1 : i2) * F + ((NR * q + role) ^ (F / 2));
    }
    best = SPeak{zmax, first};
  }
#pragma unroll
  for (int off = HW / 2; off > 0; off >>= 1) {
    int o_off = off;
    asm volatile("" : "+v"(o_off));
    SPeak other{__shfl_xor(best.v, o_off, 64), __shfl_xor(best.i, o_off, 64)};
    best = sbetter(best, other);
  }
  if (n == 0) red_p[g][role] = best;
  __syncthreads();
  if (t == 0 && valid) {
    best = red_p[g][0];
    for (int k = 1; k < NR; ++k) best = sbetter(best, red_p[g][k]);
    const float scale = 1.0f / (float)(4 * F * F);  // 4: the A / B split's factors of 2
    const int ip = best.i / F;
    const int jp = best.i % F;
    float isp = (float)ip, jsp = (float)jp;
    if (ip > 0 && ip < F - 1 && jp > 0 && jp < F - 1) {
      const float c = plane[ip * SP + jp] * scale;
      const float cd = plane[(ip - 1) * SP + jp] * scale;
      const float cu = plane[(ip + 1) * SP + jp] * scale;
      const float cl = plane[ip * SP + jp - 1] * scale;
      const float cr = plane[ip * SP + jp + 1] * scale;
      const float eps = 1e-7f;
      if (a.subpixel == kGaussian) {
        if (c > 0.f && cd > 0.f && cu > 0.f && cl > 0.f && cr > 0.f) {
          const float lc = logf(c), lcd = logf(cd), lcu = logf(cu), lcl = logf(cl), lcr = logf(cr);
          isp = (float)ip + 0.5f * (lcd - lcu) / (lcd - 2.0f * lc + lcu + eps);
          jsp = (float)jp + 0.5f * (lcl - lcr) / (lcl - 2.0f * lc + lcr + eps);
        }
      } else if (a.subpixel == kParabolic) {
        isp = (float)ip + 0.5f * (cd - cu) / (cd - 2.0f * c + cu + eps);
        jsp = (float)jp + 0.5f * (cl - cr) / (cl - 2.0f * c + cr + eps);
      } else {
        isp = (float)ip + (cu - cd) / (cd + c + cu + eps);
        jsp = (float)jp + (cr - cl) / (cl + c + cr + eps);
      }
    }
    float uu = jsp - (float)(F / 2);
    float vv = isp - (float)(F / 2);
    if (a.up) {
      uu += pred_u;
      vv += pred_v;
    }
    a.u[win] = uu;
    a.v[win] = vv;
    if (a.i_peak) {
      a.i_peak[win] = ip;
      a.j_peak[win] = jp;
    }
  }
  if (a.planes && valid) {  // uniform over the launch
    const float scale = 1.0f / (float)(4 * F * F);
    float* gp = a.planes + (size_t)win * F * F;
    for (int idx = t; idx < F * F; idx += 2 * F) gp[idx] = plane[(idx / F) * SP + idx % F] * scale;
  }
}

namespace {

template <int F, typename InT, bool BATCHED>
hipError_t split_lds_attr(size_t lds) {
  static thread_local size_t attr_bytes = 0;
  if (lds > 65536 && lds > attr_bytes) {
    hipError_t e = hipFuncSetAttribute((const void*)corr_split_kernel<F, InT, BATCHED>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) {
      set_launch_detail("hipFuncSetAttribute(MaxDynamicSharedMemorySize)", e);
      return e;
    }
    attr_bytes = lds;
  }
  return hipSuccess;
}

template <int F, typename InT>
hipError_t launch_split_t(const CorrArgs& a, hipStream_t stream) {
  using Ly = SplitLayout<F>;
  const int n_windows = a.m * a.n;
  const size_t lds = correlate_split_lds_bytes(F);
  const int blocks = (n_windows + Ly::G - 1) / Ly::G;
  (void)hipGetLastError();
  hipError_t e = split_lds_attr<F, InT, false>(lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((corr_split_kernel<F, InT, false>), dim3(blocks), dim3(256), lds, stream, a, CorrBatch{},
                     n_windows);
  e = hipGetLastError();
  if (e != hipSuccess) set_launch_detail("corr_split_kernel launch", e);
  return e;
}

template <int F, typename InT>
hipError_t launch_split_batch_t(const CorrBatch& b, int total_blocks, hipStream_t stream) {
  const size_t lds = correlate_split_lds_bytes(F);
  (void)hipGetLastError();
  hipError_t e = split_lds_attr<F, InT, true>(lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((corr_split_kernel<F, InT, true>), dim3(total_blocks), dim3(256), lds, stream, CorrArgs{}, b, 0);
  e = hipGetLastError();
  if (e != hipSuccess) set_launch_detail("corr_split_kernel (batched) launch", e);
  return e;
}

template <int F>
hipError_t launch_split_f(const CorrArgs& a, hipStream_t stream) {
  switch (a.in_dtype) {
    case kU8: return launch_split_t<F, uint8_t>(a, stream);
    case kU16: return launch_split_t<F, uint16_t>(a, stream);
    case kF32: return launch_split_t<F, float>(a, stream);
  }
  return hipErrorInvalidValue;
}

template <int F>
hipError_t launch_split_batch_f(const CorrBatch& b, int in_dtype, int total_blocks, hipStream_t stream) {
  switch (in_dtype) {
    case kU8: return launch_split_batch_t<F, uint8_t>(b, total_blocks, stream);
    case kU16: return launch_split_batch_t<F, uint16_t>(b, total_blocks, stream);
    case kF32: return launch_split_batch_t<F, float>(b, total_blocks, stream);
  }
  return hipErrorInvalidValue;
}

}  // namespace

bool correlate_split_enabled() {
  static const bool on = [] {
    const char* e = getenv("OPENPIV_HIP_CORR_SPLIT");
    return !(e && e[0] == '0');
  }();
  return on;
}

int correlate_split_windows_per_block(int fsize) { return fsize == 128 ? SplitLayout<128>::G : SplitLayout<64>::G; }

bool correlate_split_applies(const CorrArgs& a) {
  return (a.fsize == 64 || a.fsize == 128) && 2 * a.ss <= a.fsize && a.s2n == nullptr && a.win_pad == 0 &&
         (a.in_dtype == kU8 || a.in_dtype == kU16 || a.in_dtype == kF32);
}

hipError_t launch_correlate_split(const CorrArgs& a, hipStream_t stream) {
  switch (a.fsize) {
    case 64: return launch_split_f<64>(a, stream);
    case 128: return launch_split_f<128>(a, stream);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_correlate_split_batch(const CorrBatch& b, int fsize, int in_dtype, int total_blocks,
                                        hipStream_t stream) {
  if (total_blocks <= 0) return hipSuccess;
  switch (fsize) {
    case 64: return launch_split_batch_f<64>(b, in_dtype, total_blocks, stream);
    case 128: return launch_split_batch_f<128>(b, in_dtype, total_blocks, stream);
  }
  return hipErrorInvalidValue;
}

}  // namespace piv
