// Unweighted DCT-PLS smoothing with a GCV-chosen smoothing parameter
// (SmoothnGPU.__call__, gpu_smoothn.py:123-231, 299-372), for the two
// velocity fields of one pass.
//
//   z = IDCT2( Gamma(p) * DCT2(f) ),  Gamma = 1 / (1 + 10^p Lambda)
//
// The orthonormal 2-D DCT-II and its inverse run as four small tiled FP32
// GEMMs against precomputed cosine matrices.  The GCV score
//   g(p) = RSS(p) / N / (1 - T(p)/N)^2 / n_fields,
//   RSS = sum_fields ||F (Gamma - 1)||^2,  T = sum Gamma      (:353-372)
// and the search for p run in FP64 on the device: the reference's 10-point
// span over [p_min, p_max] (:182-188), then, instead of host L-BFGS-B
// (:191-193), two 64-point bracket refinements of dg/dp = 0 and two Newton
// steps using the analytic first and second derivatives of RSS and T.
#include <math.h>

#include "piv_kernels.h"

namespace piv {

namespace {

// ---------------------------------------------------------------------------
// Small batched GEMM for the 2-D DCTs: for f in {0, 1}
//   C_f(r, c) = sum_k A_f(r, k) B_f(k, c),  A(r,k) = ta ? A[k*lda + r] : A[r*lda + k]
//                                           B(k,c) = tb ? B[c*ldb + k] : B[k*ldb + c]
// 16 x 16 output tile per 256-thread workgroup (one output per thread), K
// staged through LDS in chunks of 16; FP32 with FMA accumulation.
// ---------------------------------------------------------------------------
struct GemmArgs {
  const float* a[2];
  const float* b[2];
  float* c[2];
  int lda, ldb, ldc;
  int ta, tb;
  int M, N, K;
};

__global__ __launch_bounds__(256) void gemm_small_kernel(GemmArgs g) {
  __shared__ float As[16][17];
  __shared__ float Bs[16][17];
  const int f = blockIdx.z;
  const float* A = g.a[f];
  const float* B = g.b[f];
  const int r0 = blockIdx.y * 16, c0 = blockIdx.x * 16;
  const int ty = threadIdx.x / 16, tx = threadIdx.x % 16;
  float acc = 0.f;
  for (int k0 = 0; k0 < g.K; k0 += 16) {
    {
      // As[row][k]: the thread index runs along the contiguous direction
      const int rr = g.ta ? tx : ty, kk = g.ta ? ty : tx;
      const int r = r0 + rr, k = k0 + kk;
      float val = 0.f;
      if (r < g.M && k < g.K) val = g.ta ? A[(size_t)k * g.lda + r] : A[(size_t)r * g.lda + k];
      As[rr][kk] = val;
    }
    {
      const int kk = g.tb ? tx : ty, cc = g.tb ? ty : tx;
      const int k = k0 + kk, c = c0 + cc;
      float val = 0.f;
      if (k < g.K && c < g.N) val = g.tb ? B[(size_t)c * g.ldb + k] : B[(size_t)k * g.ldb + c];
      Bs[kk][cc] = val;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < 16; ++kk) acc = fmaf(As[ty][kk], Bs[kk][tx], acc);
    __syncthreads();
  }
  const int r = r0 + ty, c = c0 + tx;
  if (r < g.M && c < g.N) g.c[f][(size_t)r * g.ldc + c] = acc;
}

hipError_t gemm2(const float* a0, const float* a1, int lda, int ta, const float* b0, const float* b1, int ldb, int tb,
                 float* c0, float* c1, int ldc, int M, int N, int K, hipStream_t s) {
  GemmArgs g{{a0, a1}, {b0, b1}, {c0, c1}, lda, ldb, ldc, ta, tb, M, N, K};
  dim3 grid((N + 15) / 16, (M + 15) / 16, 2);
  hipLaunchKernelGGL(gemm_small_kernel, grid, dim3(256), 0, s, g);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// GCV search.  Every evaluation launch spreads the m*n spectrum over
// kGcvBlocks workgroups; thread t of a block evaluates p-value (t % 64) over
// a quarter of the block's elements, so one launch scores up to 64 p-values.
// A one-block decide launch then reduces the per-block partials (stored
// p-major so each sum reads one contiguous run, summed in block order:
// deterministic) and picks the next p-values:
//   phase 0: the 10-point span of the reference, first argmin -> bracket
//   phase 1, 2: 64 evenly spaced points in the bracket, pick the sign change
//            of dg/dp nearest the span minimum -> 63x narrower bracket
//   phase 3, 4: safeguarded Newton steps on dg/dp = 0
// then a separate launch scales the spectrum by Gamma(p).
// ---------------------------------------------------------------------------
constexpr int kGcvBlocks = 128;  // the decide kernel reduces 8 lanes x 16 partials
constexpr int kSpan = 10;  // N_P0, gpu_smoothn.py:33
constexpr int kGrid = 64;

struct GcvState {
  double pv[kGrid];     // p-values of the pending evaluation
  int npv;
  double a, b;          // bracket
  double p_best_span;
  int done;             // minimum fixed at a bound / by the grid
  double p;             // current estimate
  double p_min, p_max;
};

// Sums per p-value: RSS, T, RSS', T', RSS'', T'' (L = ln 10, e = 1 - Gamma):
//   RSS' = 2L sum S e^2 G ; RSS'' = 2L^2 sum S e^2 G (3G - 1)
//   T'   = -L sum G e    ; T''   = L^2 sum G e (1 - 2G)
__device__ __forceinline__ void gcv_terms(double S, double lam, double s, double (&t)[6]) {
  const double G = 1.0 / (1.0 + s * lam);
  const double e = 1.0 - G;
  const double se2 = S * e * e;
  t[0] = se2;
  t[1] = G;
  t[2] = se2 * G;
  t[3] = G * e;
  t[4] = se2 * G * (3.0 * G - 1.0);
  t[5] = G * e * (1.0 - 2.0 * G);
}

__device__ double phi_of(const double* sm, int N) {
  const double L = 2.302585092994045684;
  const double rss = sm[0], T = sm[1], drss = 2.0 * L * sm[2], dT = -L * sm[3];
  return drss * (N - T) + 2.0 * rss * dT;
}
__device__ double dphi_of(const double* sm, int N) {
  const double L = 2.302585092994045684;
  const double rss = sm[0], T = sm[1], drss = 2.0 * L * sm[2], dT = -L * sm[3];
  const double d2rss = 2.0 * L * L * sm[4], d2T = L * L * sm[5];
  return d2rss * (N - T) + drss * dT + 2.0 * rss * d2T;
}
__device__ double g_of(const double* sm, int N) {
  const double d = 1.0 - sm[1] / N;
  return sm[0] / N / (d * d) / 2.0;
}

// One evaluation: every block scores the pending p-values over its slice of
// the spectrum and writes per-block partials (p-major).  No cross-block
// communication inside the kernel; the decide kernel runs after it.
template <int NS>
__global__ __launch_bounds__(256) void gcv_eval_kernel(const float* Y, const double* lam, int N, const GcvState* st,
                                                       double* part, int span, double p_min, double p_max) {
  __shared__ double red[256][NS];
  const int tid = threadIdx.x;
  const int pi = tid % 64, sub = tid / 64;
  // the first evaluation scores the reference's 10-point span directly
  const int npv = span ? kSpan : st->npv;
  const bool done = span ? false : (st->done != 0);
  double acc[NS];
#pragma unroll
  for (int q = 0; q < NS; ++q) acc[q] = 0.0;
  if (pi < npv && !done) {
    const double pv = span ? (double)pi * (1.0 / (kSpan - 1)) * (p_max - p_min) + p_min : st->pv[pi];
    const double s = pow(10.0, pv);
    const int chunk = (N + kGcvBlocks - 1) / kGcvBlocks;
    const int k0 = blockIdx.x * chunk, k1 = min(N, k0 + chunk);
    // four independent elements per step keep several loads in flight
    for (int k = k0 + sub; k < k1; k += 16) {
      double S4[4], L4[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int kk = k + 4 * e;
        const bool in = kk < k1;
        const double y0 = in ? (double)Y[kk] : 0.0, y1 = in ? (double)Y[N + kk] : 0.0;
        S4[e] = y0 * y0 + y1 * y1;
        L4[e] = in ? lam[kk] : 0.0;
      }
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        double t[6];
        gcv_terms(S4[e], L4[e], s, t);
        const double w = (k + 4 * e < k1) ? 1.0 : 0.0;
#pragma unroll
        for (int q = 0; q < NS; ++q) acc[q] += w * t[q];
      }
    }
  }
#pragma unroll
  for (int q = 0; q < NS; ++q) red[tid][q] = acc[q];
  __syncthreads();
  if (tid < 64) {
#pragma unroll
    for (int q = 0; q < NS; ++q) {
      const double v = red[tid][q] + red[tid + 64][q] + red[tid + 128][q] + red[tid + 192][q];
      part[((size_t)tid * 6 + q) * kGcvBlocks + blockIdx.x] = v;
    }
  }
}

// Reduces the partials of the last evaluation (block order: deterministic)
// and picks the next p-values (phases as described above).
template <int PHASE>
__global__ __launch_bounds__(1024) void gcv_decide_kernel(int N, GcvState* st, const double* part, double p_min,
                                                          double p_max) {
  constexpr int NS = (PHASE >= 3) ? 6 : 4;
  __shared__ double tot[kGrid][6];
  const int tid = threadIdx.x;
  if (PHASE == 0 && tid == 0) {
    st->npv = kSpan;
    for (int i = 0; i < kSpan; ++i) st->pv[i] = (double)i * (1.0 / (kSpan - 1)) * (p_max - p_min) + p_min;
    st->p_min = p_min;
    st->p_max = p_max;
    st->done = 0;
  }
  for (int j = tid; j < kGrid * 6; j += 1024) tot[j / 6][j % 6] = 0.0;
  __syncthreads();
  const int npv = st->npv;
  // teams of 8 consecutive lanes reduce one (p-value, sum) row of 128
  // partials: 16 per lane, all loads in flight, then a 3-step shuffle
  const int team = tid >> 3, mate = tid & 7;
  const int rows = npv * NS;
  for (int r = team; r < rows; r += 128) {
    const int pj = r / NS, q = r % NS;
    const double2* src = reinterpret_cast<const double2*>(part + ((size_t)pj * 6 + q) * kGcvBlocks) + mate * 8;
    double2 x[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) x[e] = src[e];
    double t = 0.0;
#pragma unroll
    for (int e = 0; e < 8; ++e) t += x[e].x + x[e].y;
    t += __shfl_xor(t, 4, 64);
    t += __shfl_xor(t, 2, 64);
    t += __shfl_xor(t, 1, 64);
    if (mate == 0) tot[pj][q] = t;
  }
  __syncthreads();
  if (tid >= 64) return;
  // wave 0 decides; lane i owns p-value i
  const int lane = tid;
  const bool done = st->done != 0;
  const double pvi = lane < npv ? st->pv[lane] : 0.0;
  const double gi = lane < npv ? g_of(tot[lane], N) : 1e300;
  const double fi = lane < npv ? phi_of(tot[lane], N) : 0.0;
  const double fnext = __shfl_down(fi, 1, 64);
  const double pnext = __shfl_down(pvi, 1, 64);
  double a = st->a, b = st->b, pnew = st->p;
  int newly_done = 0;
  if (!done) {
    if (PHASE == 0) {
      // first argmin of the span scores (ties -> lowest index)
      double gbest = gi;
      int ibest = lane;
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        const double go = __shfl_xor(gbest, off, 64);
        const int io = __shfl_xor(ibest, off, 64);
        if (go < gbest || (go == gbest && io < ibest)) {
          gbest = go;
          ibest = io;
        }
      }
      const double pb = __shfl(pvi, ibest, 64);
      const double pl = __shfl(pvi, ibest > 0 ? ibest - 1 : 0, 64);
      const double pr = __shfl(pvi, ibest < kSpan - 1 ? ibest + 1 : 0, 64);
      if (lane == 0) st->p_best_span = pb;
      pnew = pb;
      a = ibest > 0 ? pl : st->p_min;
      b = ibest < kSpan - 1 ? pr : st->p_max;
    } else if (PHASE <= 2) {
      // a minimum sits where phi goes - -> +; take the one nearest the span
      // minimum (ties -> lowest index)
      const double pbs = st->p_best_span;
      const bool sc = lane + 1 < npv && fi < 0.0 && fnext >= 0.0;
      double dist = sc ? fabs(0.5 * (pvi + pnext) - pbs) : 1e300;
      int jbest = sc ? lane : 1 << 20;
      double gbest = gi;
      int ibest = lane;
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        const double dO = __shfl_xor(dist, off, 64);
        const int jO = __shfl_xor(jbest, off, 64);
        if (dO < dist || (dO == dist && jO < jbest)) {
          dist = dO;
          jbest = jO;
        }
        const double go = __shfl_xor(gbest, off, 64);
        const int io = __shfl_xor(ibest, off, 64);
        if (go < gbest || (go == gbest && io < ibest)) {
          gbest = go;
          ibest = io;
        }
      }
      const double f_first = __shfl(fi, 0, 64), f_last = __shfl(fi, npv - 1, 64);
      const double p_first = __shfl(pvi, 0, 64), p_last = __shfl(pvi, npv - 1, 64);
      if (jbest < (1 << 20)) {
        const double fj = __shfl(fi, jbest, 64), fj1 = __shfl(fi, jbest + 1, 64);
        const double pj = __shfl(pvi, jbest, 64), pj1 = __shfl(pvi, jbest + 1, 64);
        a = pj;
        b = pj1;
        pnew = (fj * pj1 - fj1 * pj) / (fj - fj1);  // secant root
      } else if (f_first >= 0.0 && p_first <= st->p_min) {
        pnew = st->p_min;  // increasing from the lower bound
        newly_done = 1;
      } else if (f_last <= 0.0 && p_last >= st->p_max) {
        pnew = st->p_max;  // decreasing up to the upper bound
        newly_done = 1;
      } else {
        pnew = __shfl(pvi, ibest, 64);
        newly_done = 1;
      }
    } else {
      // Newton on phi, kept inside the bracket
      const double f = __shfl(fi, 0, 64), df = dphi_of(tot[0], N);
      const double p0 = __shfl(pvi, 0, 64);
      if (f < 0.0) a = p0; else b = p0;
      double pn = (df > 0.0) ? p0 - f / df : 0.5 * (a + b);
      if (!(pn >= a && pn <= b)) pn = 0.5 * (a + b);
      pnew = pn;
    }
  }
  // next evaluation points, one lane each
  if (PHASE == 0 || PHASE == 1) {
    st->pv[lane] = a + (b - a) * (double)lane / (kGrid - 1);
  } else if (lane == 0) {
    st->pv[0] = pnew;
  }
  if (lane == 0) {
    st->a = a;
    st->b = b;
    st->p = pnew;
    st->npv = (PHASE == 0 || PHASE == 1) ? kGrid : 1;
    if (newly_done) st->done = 1;
  }
}

// z_dct = Gamma * f_dct (gpu_smoothn.py:196-199); p from the search state
// (p_state) or fixed (p_state == null).
__global__ void gamma_apply_kernel(float* Y, const double* lam, int N, double p_fixed, const GcvState* p_state,
                                   double* p_out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const double p = p_state ? p_state->p : p_fixed;
  if (k < N) {
    const double G = 1.0 / (1.0 + pow(10.0, p) * lam[k]);
    Y[k] = (float)(G * (double)Y[k]);
    Y[N + k] = (float)(G * (double)Y[N + k]);
  }
  if (k == 0 && p_out) *p_out = p;
}

}  // namespace

constexpr size_t kGcvPartOffset = (sizeof(GcvState) + 255) & ~size_t(255);
size_t smoothn_state_bytes() { return kGcvPartOffset + (size_t)kGcvBlocks * kGrid * 6 * sizeof(double); }

hipError_t launch_smoothn(float* u, float* v, int m, int n, double p_min, double p_max, double fixed_p, int auto_s,
                          const SmoothnWork& w, hipStream_t stream) {
  const int N = m * n;
  if (N < 2) return hipSuccess;  // gpu_smoothn.py:137-140
  hipError_t e;
  // forward orthonormal DCT-II: T0 = X Cn^T ; T1 = Cm T0
  e = gemm2(u, v, n, 0, w.cos_n, w.cos_n, n, 1, w.t0, w.t0 + N, n, m, n, n, stream);
  if (e != hipSuccess) return e;
  e = gemm2(w.cos_m, w.cos_m, m, 0, w.t0, w.t0 + N, n, 0, w.t1, w.t1 + N, n, m, n, m, stream);
  if (e != hipSuccess) return e;
  if (auto_s) {
    GcvState* st = reinterpret_cast<GcvState*>(w.state);
    double* part = reinterpret_cast<double*>(reinterpret_cast<char*>(w.state) + kGcvPartOffset);
    const dim3 eb(kGcvBlocks), et(256);
    hipLaunchKernelGGL(gcv_eval_kernel<4>, eb, et, 0, stream, w.t1, w.lambda, N, st, part, 1, p_min, p_max);
    hipLaunchKernelGGL(gcv_decide_kernel<0>, dim3(1), dim3(1024), 0, stream, N, st, part, p_min, p_max);
    hipLaunchKernelGGL(gcv_eval_kernel<4>, eb, et, 0, stream, w.t1, w.lambda, N, st, part, 0, p_min, p_max);
    hipLaunchKernelGGL(gcv_decide_kernel<1>, dim3(1), dim3(1024), 0, stream, N, st, part, p_min, p_max);
    hipLaunchKernelGGL(gcv_eval_kernel<4>, eb, et, 0, stream, w.t1, w.lambda, N, st, part, 0, p_min, p_max);
    hipLaunchKernelGGL(gcv_decide_kernel<2>, dim3(1), dim3(1024), 0, stream, N, st, part, p_min, p_max);
    hipLaunchKernelGGL(gcv_eval_kernel<6>, eb, et, 0, stream, w.t1, w.lambda, N, st, part, 0, p_min, p_max);
    hipLaunchKernelGGL(gcv_decide_kernel<3>, dim3(1), dim3(1024), 0, stream, N, st, part, p_min, p_max);
    hipLaunchKernelGGL(gcv_eval_kernel<6>, eb, et, 0, stream, w.t1, w.lambda, N, st, part, 0, p_min, p_max);
    hipLaunchKernelGGL(gcv_decide_kernel<4>, dim3(1), dim3(1024), 0, stream, N, st, part, p_min, p_max);
    hipLaunchKernelGGL(gamma_apply_kernel, dim3((N + 255) / 256), dim3(256), 0, stream, w.t1, w.lambda, N, 0.0,
                       (const GcvState*)st, w.p_out);
  } else {
    hipLaunchKernelGGL(gamma_apply_kernel, dim3((N + 255) / 256), dim3(256), 0, stream, w.t1, w.lambda, N, fixed_p,
                       (const GcvState*)nullptr, w.p_out);
  }
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  // inverse: T0 = Cm^T T1 ; Z = T0 Cn
  e = gemm2(w.cos_m, w.cos_m, m, 1, w.t1, w.t1 + N, n, 0, w.t0, w.t0 + N, n, m, n, m, stream);
  if (e != hipSuccess) return e;
  return gemm2(w.t0, w.t0 + N, n, 0, w.cos_n, w.cos_n, n, 0, u, v, n, m, n, n, stream);
}

}  // namespace piv
