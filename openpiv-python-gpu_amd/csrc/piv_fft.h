// Register-resident radix-2 FFT building blocks for gfx950 (wave64).
//
// A length-F 1-D transform is spread over 2 or 4 adjacent lanes; each lane
// keeps F/2 or F/4 complex values in VGPRs, runs an in-register DIF FFT on
// them (fully unrolled, compile-time twiddles) and exchanges data with its
// partner lanes through DPP quad permutes -- no LDS traffic inside a 1-D
// transform.  Register placement after an in-register DIF is bit-reversed;
// all indices are compile-time constants, so the permutation costs nothing.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace piv {

// Compile-time loop: fn(std::integral_constant<int, I>) for I in [B, E).
template <int B, int E, typename Fn>
__device__ __forceinline__ void static_for(Fn&& fn) {
  if constexpr (B < E) {
    fn(std::integral_constant<int, B>{});
    static_for<B + 1, E>(fn);
  }
}

// ---------------------------------------------------------------------------
// Compile-time trigonometry (double precision Taylor series after octant
// reduction; the float casts are correctly rounded for every twiddle used).
// ---------------------------------------------------------------------------
constexpr double kPi = 3.14159265358979323846264338327950288;

constexpr double taylor_sin(double x) {
  double term = x, sum = x;
  for (int k = 1; k < 14; ++k) {
    term *= -x * x / ((2.0 * k) * (2.0 * k + 1.0));
    sum += term;
  }
  return sum;
}
constexpr double taylor_cos(double x) {
  double term = 1.0, sum = 1.0;
  for (int k = 1; k < 14; ++k) {
    term *= -x * x / ((2.0 * k - 1.0) * (2.0 * k));
    sum += term;
  }
  return sum;
}
// cos(2*pi*k/n) and sin(2*pi*k/n) for integers k, n (exact rational
// reduction to the first octant, then a Taylor series).
constexpr double cos_frac(long k, long n) {
  k %= n;
  if (k < 0) k += n;
  if (k == 0) return 1.0;
  if (2 * k > n) k = n - k;                           // cos(2pi - a) = cos(a)
  if (4 * k == n) return 0.0;
  if (4 * k > n) return -cos_frac(n - 2 * k, 2 * n);  // cos(a) = -cos(pi - a)
  if (8 * k > n) return taylor_sin(2.0 * kPi * (double)(n - 4 * k) / (4.0 * (double)n));
  return taylor_cos(2.0 * kPi * (double)k / (double)n);
}
constexpr double sin_frac(long k, long n) { return cos_frac(4 * k - n, 4 * n); }

struct c32 {
  float x, y;
};

__device__ __forceinline__ c32 cadd(c32 a, c32 b) { return {a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ c32 csub(c32 a, c32 b) { return {a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ c32 cmul(c32 a, c32 b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// conj(a) * b
__device__ __forceinline__ c32 cconjmul(c32 a, c32 b) {
  return {a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x};
}

// Multiply by exp(DIR * 2*pi*i * K / N) with the trivial cases folded.
template <int K, int N, int DIR>
__device__ __forceinline__ c32 twiddle(c32 v) {
  constexpr int k = ((K % N) + N) % N;
  if constexpr (k == 0) {
    return v;
  } else if constexpr (4 * k == N) {  // DIR * i
    return DIR > 0 ? c32{-v.y, v.x} : c32{v.y, -v.x};
  } else if constexpr (2 * k == N) {
    return c32{-v.x, -v.y};
  } else if constexpr (4 * k == 3 * N) {  // -DIR * i
    return DIR > 0 ? c32{v.y, -v.x} : c32{-v.y, v.x};
  } else {
    constexpr float c = (float)cos_frac(k, N);
    constexpr float s = (float)(DIR * sin_frac(k, N));
    return c32{v.x * c - v.y * s, v.x * s + v.y * c};
  }
}

constexpr int ilog2(int n) { return n <= 1 ? 0 : 1 + ilog2(n / 2); }
constexpr int bitrev(int v, int n) {
  int bits = ilog2(n), r = 0;
  for (int i = 0; i < bits; ++i) r |= ((v >> i) & 1) << (bits - 1 - i);
  return r;
}

// ---------------------------------------------------------------------------
// In-register FFT of length L: natural-order input x[n]; on return the
// frequency-k value sits in x[bitrev(k, L)].
//
// Radix-2 decimation in time on the bit-reversed view y[i] = x[bitrev(i)]:
// DIT takes its input bit-reversed and returns its output in natural order,
// so with this view both conventions above hold and the permutation is only
// register naming.  A butterfly with a non-trivial twiddle w is
//   u' = u + w v  (four FMAs),   v' = 2u - u'  (two FMAs),
// six instructions instead of the eight of a multiply and two add pairs;
// twiddles 1, -1, +-i cost only the adds.  HALF_ZERO: inputs x[L/2..L-1]
// are zero, so the first stage (pairs x[n], x[n + L/2]) only copies.
// ---------------------------------------------------------------------------
template <int K, int N, int DIR>
__device__ __forceinline__ void dit_bfly(c32& a, c32& b) {
  constexpr int k = ((K % N) + N) % N;
  if constexpr (k == 0 || 4 * k == N || 2 * k == N || 4 * k == 3 * N) {
    const c32 t = twiddle<K, N, DIR>(b);
    const c32 u = a;
    a = cadd(u, t);
    b = csub(u, t);
  } else {
    constexpr float c = (float)cos_frac(k, N);
    constexpr float s = (float)(DIR * sin_frac(k, N));
    const c32 u = a, v = b;
    const float ux = fmaf(v.x, c, fmaf(-v.y, s, u.x));
    const float uy = fmaf(v.x, s, fmaf(v.y, c, u.y));
    a = c32{ux, uy};
    b = c32{fmaf(2.0f, u.x, -ux), fmaf(2.0f, u.y, -uy)};
  }
}

template <int L, int DIR, bool HALF_ZERO = false>
__device__ __forceinline__ void fft_reg(c32 (&x)[L]) {
  static_for<0, ilog2(L)>([&](auto SS) {
    constexpr int half = 1 << decltype(SS)::value, m = 2 * half;
    if constexpr (HALF_ZERO && half == 1) {
      // y[2q + 1] = x[bitrev(2q) + L/2] = 0: the butterfly copies y[2q]
      static_for<0, L / 2>([&](auto QQ) {
        constexpr int q = decltype(QQ)::value;
        x[bitrev(2 * q + 1, L)] = x[bitrev(2 * q, L)];
      });
    } else {
      static_for<0, L / 2>([&](auto BB) {
        constexpr int bi = decltype(BB)::value;  // butterfly index
        constexpr int blk = bi / half, j = bi % half;
        constexpr int i0 = blk * m + j, i1 = i0 + half;
        dit_bfly<j * (L / m), L, DIR>(x[bitrev(i0, L)], x[bitrev(i1, L)]);
      });
    }
  });
}

// ---------------------------------------------------------------------------
// Cross-lane moves (DPP quad permutes; all four lanes of a quad must be
// active).  bound_ctrl is set: with every source lane valid it changes
// nothing, and it lets the compiler fold the move into the consuming add.
// ---------------------------------------------------------------------------
template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), CTRL, 0xF, 0xF, true));
}
template <int CTRL>
__device__ __forceinline__ c32 dpp_c(c32 v) {
  return c32{dpp_f<CTRL>(v.x), dpp_f<CTRL>(v.y)};
}
constexpr int kSwap1 = 0xB1;    // quad_perm [1,0,3,2]: lane ^ 1
constexpr int kSwap2 = 0x4E;    // quad_perm [2,3,0,1]: lane ^ 2
constexpr int kReverse = 0x1B;  // quad_perm [3,2,1,0]: lane ^ 3

__device__ __forceinline__ c32 swap1(c32 v) { return dpp_c<kSwap1>(v); }
__device__ __forceinline__ c32 swap2(c32 v) { return dpp_c<kSwap2>(v); }

// ---------------------------------------------------------------------------
// Cross-lane radix-2 butterflies.  The lane that holds the odd half
// pre-multiplies it by the twiddle, both lanes swap, and each lane finishes
// with one signed add:  lower: own + recv,  upper: recv - own_twiddled.
// ---------------------------------------------------------------------------
template <int CTRL>
__device__ __forceinline__ c32 xlane(c32 v) { return dpp_c<CTRL>(v); }

// Before: lane h (h = upper) holds the length-F/2 transform of the samples
// x[2m + h] (frequency k at register bitrev(k, F/2)).  After: lane h holds
// X[k + h*F/2] at register bitrev(k, F/2).
template <int F, int DIR, int CTRL = kSwap1>
__device__ __forceinline__ void dit2_combine(c32 (&x)[F / 2], bool upper) {
  static_for<0, F / 2>([&](auto KK) {
    constexpr int K = decltype(KK)::value;
    constexpr int r = bitrev(K, F / 2);
    const c32 own = x[r];
    const c32 tw = twiddle<K, F, DIR>(own);
    const c32 v = upper ? tw : own;
    const c32 recv = xlane<CTRL>(v);
    const c32 sv = upper ? c32{-v.x, -v.y} : v;
    x[r] = c32{recv.x + sv.x, recv.y + sv.y};
  });
}

// 2-lane DIF first stage.  Before: lane h holds x[n + h*F/2] at register
// bitrev(n, F/2).  After: lane 0 holds y[n] = x[n] + x[n+F/2], lane 1 holds
// y'[n] = (x[n] - x[n+F/2]) w^n, at the same registers.
template <int F, int DIR, int CTRL = kSwap1>
__device__ __forceinline__ void dif2_split(c32 (&x)[F / 2], bool upper) {
  static_for<0, F / 2>([&](auto NN) {
    constexpr int N = decltype(NN)::value;
    constexpr int r = bitrev(N, F / 2);
    const c32 own = x[r];
    const c32 recv = xlane<CTRL>(own);
    // lower: x[n] + x[n+H];  upper: x[n] - x[n+H] = recv - own
    const c32 t = upper ? c32{recv.x - own.x, recv.y - own.y} : c32{own.x + recv.x, own.y + recv.y};
    x[r] = upper ? twiddle<N, F, DIR>(t) : t;
  });
}

}  // namespace piv
