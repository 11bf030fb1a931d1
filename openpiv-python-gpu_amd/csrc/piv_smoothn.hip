// Unweighted DCT-PLS smoothing with a GCV-chosen smoothing parameter
// (SmoothnGPU.__call__, gpu_smoothn.py:123-231, 299-372), for the two
// velocity fields of one pass.
//
//   z = IDCT2( Gamma(p) * DCT2(f) ),  Gamma = 1 / (1 + 10^p Lambda)
//
// The orthonormal 2-D DCT-II and its inverse run as four small FP32
// matrix-product kernels against precomputed cosine matrices.  The GCV score
//   g(p) = RSS(p) / N / (1 - T(p)/N)^2 / n_fields,
//   RSS = sum_fields ||F (Gamma - 1)||^2,  T = sum Gamma      (:353-372)
// and the search for p run in FP64 inside one workgroup: the reference's
// 10-point span over [p_min, p_max] (:182-188), then, instead of host
// L-BFGS-B (:191-193), a bracketed Newton/bisection on dg/dp = 0 using the
// analytic first and second derivatives of RSS and T.
#include <math.h>

#include "piv_kernels.h"

namespace piv {

namespace {

// out[f][i][l] = sum_j X[f][i][j] * M(l, j);  M(l,j) = trans ? M[j*n + l] : M[l*n + j]
__global__ __launch_bounds__(256) void apply_right_kernel(const float* x0, const float* x1, const float* M, int m,
                                                          int n, int trans, float* o0, float* o1) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  const int f = blockIdx.y;
  if (idx >= m * n) return;
  const int i = idx / n, l = idx % n;
  const float* X = (f == 0 ? x0 : x1) + (size_t)i * n;
  float acc = 0.f;
  if (trans) {
    for (int j = 0; j < n; ++j) acc = fmaf(X[j], M[(size_t)j * n + l], acc);
  } else {
    const float* Mr = M + (size_t)l * n;
    for (int j = 0; j < n; ++j) acc = fmaf(X[j], Mr[j], acc);
  }
  (f == 0 ? o0 : o1)[idx] = acc;
}

// out[f][k][l] = sum_i M(k, i) * X[f][i][l];  M(k,i) = trans ? M[i*m + k] : M[k*m + i]
__global__ __launch_bounds__(256) void apply_left_kernel(const float* X, const float* M, int m, int n, int trans,
                                                         float* o0, float* o1) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  const int f = blockIdx.y;
  if (idx >= m * n) return;
  const int k = idx / n, l = idx % n;
  const float* Xf = X + (size_t)f * m * n;
  float acc = 0.f;
  if (trans) {
    for (int i = 0; i < m; ++i) acc = fmaf(M[(size_t)i * m + k], Xf[(size_t)i * n + l], acc);
  } else {
    const float* Mr = M + (size_t)k * m;
    for (int i = 0; i < m; ++i) acc = fmaf(Mr[i], Xf[(size_t)i * n + l], acc);
  }
  (f == 0 ? o0 : o1)[idx] = acc;
}

constexpr int kGcvThreads = 1024;
constexpr int kSpan = 10;  // N_P0, gpu_smoothn.py:33

__device__ __forceinline__ double wsum(double x) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
  return x;
}

template <int K>
__device__ __forceinline__ void bsum(double (&v)[K], double* scratch /* 16*K */) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = wsum(v[k]);
  __syncthreads();
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) scratch[wave * K + k] = v[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) {
    double s = 0.0;
    for (int w = 0; w < kGcvThreads / 64; ++w) s += scratch[w * K + k];
    v[k] = s;
  }
}

struct Sums {
  double rss, drss, d2rss, tr, dtr, d2tr;
};

// Sums needed for g and its first two p-derivatives (L = ln 10, e = 1-Gamma):
//   RSS'  = 2L   sum S e^2 Gamma
//   RSS'' = 2L^2 sum S e^2 Gamma (3 Gamma - 1)
//   T'    = -L   sum Gamma e
//   T''   = L^2  sum Gamma e (1 - 2 Gamma)
__device__ Sums gcv_sums(const float* Y, const double* lam, int N, double p, double* scratch) {
  const double s = pow(10.0, p);
  const double L = 2.302585092994045684;
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (int k = threadIdx.x; k < N; k += kGcvThreads) {
    const double y0 = (double)Y[k], y1 = (double)Y[N + k];
    const double S = y0 * y0 + y1 * y1;
    const double G = 1.0 / (1.0 + s * lam[k]);
    const double e = 1.0 - G;
    const double se2g = S * e * e * G;
    acc[0] += S * e * e;
    acc[1] += se2g;
    acc[2] += se2g * (3.0 * G - 1.0);
    acc[3] += G;
    acc[4] += G * e;
    acc[5] += G * e * (1.0 - 2.0 * G);
  }
  bsum<6>(acc, scratch);
  Sums r;
  r.rss = acc[0];
  r.drss = 2.0 * L * acc[1];
  r.d2rss = 2.0 * L * L * acc[2];
  r.tr = acc[3];
  r.dtr = -L * acc[4];
  r.d2tr = L * L * acc[5];
  return r;
}

__global__ __launch_bounds__(kGcvThreads) void gcv_kernel(float* Y, const double* lam, int N, double p_min,
                                                          double p_max, double fixed_p, int auto_s,
                                                          double* p_out) {
  __shared__ double scratch[16 * 2 * kSpan];
  double p = fixed_p;
  if (auto_s) {
    // 10-point span, first argmin (gpu_smoothn.py:183-188)
    double ps[kSpan];
#pragma unroll
    for (int i = 0; i < kSpan; ++i) ps[i] = (double)i * (1.0 / (kSpan - 1)) * (p_max - p_min) + p_min;
    double acc[2 * kSpan];
#pragma unroll
    for (int i = 0; i < 2 * kSpan; ++i) acc[i] = 0.0;
    double sv[kSpan];
#pragma unroll
    for (int i = 0; i < kSpan; ++i) sv[i] = pow(10.0, ps[i]);
    for (int k = threadIdx.x; k < N; k += kGcvThreads) {
      const double y0 = (double)Y[k], y1 = (double)Y[N + k];
      const double S = y0 * y0 + y1 * y1;
      const double lk = lam[k];
#pragma unroll
      for (int i = 0; i < kSpan; ++i) {
        const double G = 1.0 / (1.0 + sv[i] * lk);
        const double e = G - 1.0;
        acc[2 * i] += S * e * e;
        acc[2 * i + 1] += G;
      }
    }
    bsum<2 * kSpan>(acc, scratch);
    int best = 0;
    double gbest = 0.0;
    for (int i = 0; i < kSpan; ++i) {
      const double d = 1.0 - acc[2 * i + 1] / N;
      const double g = acc[2 * i] / N / (d * d) / 2.0;
      if (i == 0 || g < gbest) {
        gbest = g;
        best = i;
      }
    }
    // bracket around the best span point
    double a = best > 0 ? ps[best - 1] : p_min;
    double b = best < kSpan - 1 ? ps[best + 1] : p_max;
    p = ps[best];
    // phi = RSS' (N - T) + 2 RSS T'  has the sign of dg/dp.
    auto phi_of = [&](const Sums& s) { return s.drss * (N - s.tr) + 2.0 * s.rss * s.dtr; };
    __syncthreads();
    Sums sa = gcv_sums(Y, lam, N, a, scratch);
    Sums sb = gcv_sums(Y, lam, N, b, scratch);
    const double fa = phi_of(sa), fb = phi_of(sb);
    if (fa >= 0.0 && best == 0) {
      p = p_min;  // minimum on the lower bound
    } else if (fb <= 0.0 && best == kSpan - 1) {
      p = p_max;  // minimum on the upper bound
    } else if (fa < 0.0 && fb > 0.0) {
      for (int it = 0; it < 100; ++it) {
        Sums s = gcv_sums(Y, lam, N, p, scratch);
        const double f = phi_of(s);
        const double df = s.d2rss * (N - s.tr) + s.drss * s.dtr + 2.0 * s.rss * s.d2tr;
        if (f < 0.0)
          a = p;
        else
          b = p;
        double pn = (df > 0.0) ? p - f / df : 0.5 * (a + b);
        if (!(pn > a && pn < b)) pn = 0.5 * (a + b);
        const double step = fabs(pn - p);
        p = pn;
        if (f == 0.0 || step < 1e-14 * fmax(1.0, fabs(p)) || (b - a) < 1e-14 * fmax(1.0, fabs(p))) break;
      }
    } else {
      // no sign change: golden-section search on g inside [a, b]
      const double gr = 0.6180339887498949;
      double c = b - gr * (b - a), d = a + gr * (b - a);
      auto g_of = [&](double pp) {
        Sums s = gcv_sums(Y, lam, N, pp, scratch);
        const double dd = 1.0 - s.tr / N;
        return s.rss / N / (dd * dd) / 2.0;
      };
      double gc = g_of(c), gd = g_of(d);
      for (int it = 0; it < 80 && (b - a) > 1e-12; ++it) {
        if (gc < gd) {
          b = d;
          d = c;
          gd = gc;
          c = b - gr * (b - a);
          gc = g_of(c);
        } else {
          a = c;
          c = d;
          gc = gd;
          d = a + gr * (b - a);
          gd = g_of(d);
        }
      }
      p = 0.5 * (a + b);
      const double gm = g_of(p);
      if (gbest < gm) p = ps[best];
    }
  }
  // z_dct = Gamma * f_dct  (gpu_smoothn.py:196-199)
  const double s = pow(10.0, p);
  for (int k = threadIdx.x; k < N; k += kGcvThreads) {
    const double G = 1.0 / (1.0 + s * lam[k]);
    Y[k] = (float)(G * (double)Y[k]);
    Y[N + k] = (float)(G * (double)Y[N + k]);
  }
  if (threadIdx.x == 0 && p_out) *p_out = p;
}

}  // namespace

hipError_t launch_smoothn(float* u, float* v, int m, int n, double p_min, double p_max, double fixed_p, int auto_s,
                          const SmoothnWork& w, hipStream_t stream) {
  const int N = m * n;
  if (N < 2) return hipSuccess;  // gpu_smoothn.py:137-140
  dim3 grid((N + 255) / 256, 2);
  // T0 = X Cn^T ; T1 = Cm T0                (forward DCT-II, orthonormal)
  hipLaunchKernelGGL(apply_right_kernel, grid, dim3(256), 0, stream, u, v, w.cos_n, m, n, 0, w.t0, w.t0 + N);
  hipLaunchKernelGGL(apply_left_kernel, grid, dim3(256), 0, stream, w.t0, w.cos_m, m, n, 0, w.t1, w.t1 + N);
  hipLaunchKernelGGL(gcv_kernel, dim3(1), dim3(kGcvThreads), 0, stream, w.t1, w.lambda, N, p_min, p_max, fixed_p,
                     auto_s, w.p_out);
  // T0 = Cm^T T1 ; Z = T0 Cn                (inverse)
  hipLaunchKernelGGL(apply_left_kernel, grid, dim3(256), 0, stream, w.t1, w.cos_m, m, n, 1, w.t0, w.t0 + N);
  hipLaunchKernelGGL(apply_right_kernel, grid, dim3(256), 0, stream, w.t0, w.t0 + N, w.cos_n, m, n, 1, u, v);
  return hipGetLastError();
}

}  // namespace piv
