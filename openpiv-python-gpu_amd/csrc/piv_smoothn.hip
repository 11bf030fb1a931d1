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
// (:191-193), a 64-point bracket refinement of dg/dp = 0, a secant root
// and two Newton steps using the analytic first and second derivatives of
// RSS and T.
#include <limits.h>
#include <math.h>

#include <algorithm>

#include "piv_kernels.h"
#include "piv_record.h"
#include "piv_wave.h"

namespace piv {

namespace {

__device__ __forceinline__ float sfma(float a, float b, float c) { return fmaf(a, b, c); }
__device__ __forceinline__ double sfma(double a, double b, double c) { return fma(a, b, c); }

// ---------------------------------------------------------------------------
// GCV search, three launches and a final decision inside the consumer:
//   grid:    kGrid p-values evenly spaced over [p_min, p_max] (every seventh
//            is a point of the reference's 10-point span), one per lane; the
//            spectrum split over up to kGcvBlocks workgroups -> partials
//   newton1: every workgroup sums the grid partials in a fixed order and
//            decides (first argmin of the span -> bracket between its
//            neighbours; in it, the sign change of dg/dp nearest the span
//            minimum, 9x finer than the span -> secant root p1), then scores
//            p1 over its part of the spectrum -> partials
//   newton2: every workgroup decides p2 (safeguarded Newton step on
//            dg/dp = 0 from p1), scores it -> partials
//   final:   the consumer (the inverse DCT that scales the spectrum by
//            Gamma(p), or a one-workgroup kernel) decides p3 the same way.
// Every workgroup of a launch reaches the same decision from the same
// partials (deterministic sums); workgroup 0 also writes it to the search
// state for the next launch.  Partials pass between launches only, so no
// workgroup waits for another and no cache maintenance is needed.
// ---------------------------------------------------------------------------
constexpr int kGcvThreads = 256;
constexpr int kGcvBlocks = 128;              // at most, grid launch (128 elements each at least)
constexpr int kGcvNewtonBlocks = 64;         // at most, per Newton step
constexpr int kSpan = 10;  // N_P0, gpu_smoothn.py:33
constexpr int kGrid = 64;
constexpr int kSpanStep = (kGrid - 1) / (kSpan - 1);  // grid points per span interval
static_assert((kGrid - 1) % (kSpan - 1) == 0, "span points must lie on the grid");

struct GcvState {
  double pv[kGrid];     // p-values of the pending evaluation
  int npv;
  double a, b;          // bracket
  double p_best_span;
  int done;             // minimum fixed at a bound / by the grid
  double p;             // current estimate
  double p_min, p_max;
  uint32_t arrivals;    // (unused by the unweighted search)
  int32_t nb[3];        // slot 3: workgroups of the grid / newton1 / newton2 launches
};

// Partials of the three launches, after the state.
constexpr size_t kGcvPart0 = 0;                                   // [b][p][4]
constexpr size_t kGcvPart1 = kGcvPart0 + (size_t)kGcvBlocks * kGrid * 4;  // [b][6]
constexpr size_t kGcvPart2 = kGcvPart1 + (size_t)kGcvNewtonBlocks * 6;    // [b][6]
constexpr size_t kGcvPartTotal = kGcvPart2 + (size_t)kGcvNewtonBlocks * 6;

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

// p-value of grid point j of the first evaluation; span points use the
// span's own formula (i / 9 of the range above p_min).
__device__ __forceinline__ double grid_p(int j, double p_min, double p_max) {
  if (j % kSpanStep == 0) return (double)(j / kSpanStep) * (1.0 / (kSpan - 1)) * (p_max - p_min) + p_min;
  return (double)j * (1.0 / (kGrid - 1)) * (p_max - p_min) + p_min;
}

// First (index) minimum of a wave-distributed value.
__device__ __forceinline__ void wave_argmin(double& v, int& i) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const double vo = __shfl_xor(v, off, 64);
    const int io = __shfl_xor(i, off, 64);
    if (vo < v || (vo == v && io < i)) {
      v = vo;
      i = io;
    }
  }
}

// Picks the next p-value from the sums tot[p-value][sum] of the evaluation
// that just finished (phases as described above); wave 0.  st may be a
// workgroup-local copy.
template <int PHASE>
__device__ void gcv_decide(int N, GcvState* st, double p_min, double p_max, const double (*tot)[6]) {
  const int lane = threadIdx.x;
  const int npv = PHASE == 0 ? kGrid : st->npv;
  const bool done = PHASE == 0 ? false : st->done != 0;
  const double pvi = PHASE == 0 ? grid_p(lane, p_min, p_max) : (lane < npv ? st->pv[lane] : 0.0);
  const double gi = lane < npv ? g_of(tot[lane], N) : 1e300;
  const double fi = lane < npv ? phi_of(tot[lane], N) : 0.0;
  const double fnext = __shfl_down(fi, 1, 64);
  const double pnext = __shfl_down(pvi, 1, 64);
  double a = PHASE == 0 ? p_min : st->a, b = PHASE == 0 ? p_max : st->b, pnew = PHASE == 0 ? p_min : st->p;
  int newly_done = 0;
  if (PHASE == 0) {
    // first argmin of the span scores (ties -> lowest index)
    double gs = lane % kSpanStep == 0 ? gi : 1e300;
    int is = lane;
    wave_argmin(gs, is);
    const int ib = is / kSpanStep;
    const int ja = ib > 0 ? is - kSpanStep : 0;
    const int jb = ib < kSpan - 1 ? is + kSpanStep : kGrid - 1;
    const double pbs = __shfl(pvi, is, 64);
    // a minimum sits where phi goes - -> +; take the one nearest the span
    // minimum inside the bracket (ties -> lowest index)
    const bool sc = lane >= ja && lane < jb && fi < 0.0 && fnext >= 0.0;
    double dist = sc ? fabs(0.5 * (pvi + pnext) - pbs) : 1e300;
    int jbest = sc ? lane : 1 << 20;
    wave_argmin(dist, jbest);
    double gbr = (lane >= ja && lane <= jb) ? gi : 1e300;
    int ibr = lane;
    wave_argmin(gbr, ibr);
    const double f_first = __shfl(fi, ja, 64), f_last = __shfl(fi, jb, 64);
    if (jbest < (1 << 20)) {
      const double fj = __shfl(fi, jbest, 64), fj1 = __shfl(fi, jbest + 1, 64);
      const double pj = __shfl(pvi, jbest, 64), pj1 = __shfl(pvi, jbest + 1, 64);
      a = pj;
      b = pj1;
      pnew = (fj * pj1 - fj1 * pj) / (fj - fj1);  // secant root
    } else if (f_first >= 0.0 && ja == 0) {
      pnew = p_min;  // increasing from the lower bound
      newly_done = 1;
    } else if (f_last <= 0.0 && jb == kGrid - 1) {
      pnew = p_max;  // decreasing up to the upper bound
      newly_done = 1;
    } else {
      pnew = __shfl(pvi, ibr, 64);
      newly_done = 1;
    }
  } else if (!done) {
    // Newton on phi, kept inside the bracket
    const double f = __shfl(fi, 0, 64), df = dphi_of(tot[0], N);
    const double p0 = __shfl(pvi, 0, 64);
    if (f < 0.0) a = p0; else b = p0;
    double pn = (df > 0.0) ? p0 - f / df : 0.5 * (a + b);
    if (!(pn >= a && pn <= b)) pn = 0.5 * (a + b);
    pnew = pn;
  }
  if (lane == 0) {
    st->pv[0] = pnew;
    st->a = a;
    st->b = b;
    st->p = pnew;
    st->npv = 1;
    if (PHASE == 0) {
      st->p_min = p_min;
      st->p_max = p_max;
      st->done = newly_done;
    } else if (newly_done) {
      st->done = 1;
    }
  }
}

// The search state is four GcvState slots: [1] the decision after the grid
// launch (written by newton1's workgroup 0), [2] the decision after newton1
// (by newton2's workgroup 0), [0] the final decision (by the consumer's
// workgroup 0; what every later reader of the search takes), and [3].nb the
// launch sizes.  A launch only reads slots written by earlier launches, so
// its workgroup 0 can publish while the others still read.
constexpr int kGcvSlots = 4;

// sum_{b < nb} p[b * stride] in b order, sixteen loads in flight at a time
// (the partials come from other workgroups' earlier launch: far away).
__device__ __forceinline__ double sum_strided(const double* p, int stride, int nb) {
  double t = 0.0;
  for (int b0 = 0; b0 < nb; b0 += 16) {
    double x[16];
#pragma unroll
    for (int e = 0; e < 16; ++e) x[e] = b0 + e < nb ? p[(size_t)(b0 + e) * stride] : 0.0;
#pragma unroll
    for (int e = 0; e < 16; ++e)
      if (b0 + e < nb) t += x[e];
  }
  return t;
}

// The decision after launch PHASE (0 grid, 1 newton1, 2 newton2) from that
// launch's partials, on a workgroup-local state: the previous decision is
// slot PHASE, the partials are summed in workgroup order.  Every thread
// returns after the decision.  Needs blockDim.x >= 256 for PHASE 0 (one
// thread per p-value and sum) and >= 64 otherwise.
template <int PHASE>
__device__ void gcv_decide_from(const GcvState* gst, const double* part_base, int N, double p_min, double p_max,
                                GcvState* ls, double (*tot)[6]) {
  const int tid = threadIdx.x;
  const int nb = gst[3].nb[PHASE];
  if (PHASE == 0) {
    const double* part = part_base + kGcvPart0;
    const int p = tid & 63, q = tid >> 6;
    if (tid < kGrid * 4) tot[p][q] = sum_strided(part + p * 4 + q, kGrid * 4, nb);
  } else {
    if (tid == 0) *ls = gst[PHASE];
    const double* part = part_base + (PHASE == 1 ? kGcvPart1 : kGcvPart2);
    if (tid < 6) tot[0][tid] = gst[PHASE].done ? 0.0 : sum_strided(part + tid, 6, nb);
  }
  __syncthreads();
  if (tid < 64) gcv_decide<PHASE>(N, ls, p_min, p_max, tot);
  __syncthreads();
}

// Publishes the decision after launch PHASE (one thread of workgroup 0).
template <int PHASE>
__device__ __forceinline__ void gcv_publish(GcvState* gst, const GcvState& ls) {
  gst[PHASE == 2 ? 0 : PHASE + 1] = ls;
}

// Grid evaluation: lane j of every wave scores grid p-value j; workgroup b
// covers elements [b chunk, (b + 1) chunk) of the spectrum, its four waves
// taking every fourth element (each element is read once per workgroup, by
// all lanes at the same address).  Partials part0[b][p][sum].
template <typename R>
struct GcvArgs {
  const R* Y;
  const double* lam;
  int N;
  GcvState* gst;
  double* part_base;
  double p_min, p_max;
  const int32_t* skip;
};
template <typename R>
struct GcvGridBody {
  using Args = GcvArgs<R>;
  static __device__ void run(const Args& a, VB vb);
};
template <typename R>
__global__ __launch_bounds__(kGcvThreads) void gcv_grid_kernel(GcvArgs<R> a) {
  GcvGridBody<R>::run(a, vb_native());
}
template <typename R>
__device__ void GcvGridBody<R>::run(const Args& ga, VB vb) {
  const R* Y = ga.Y;
  const double* lam = ga.lam;
  const int N = ga.N;
  GcvState* gst = ga.gst;
  double* part_base = ga.part_base;
  const double p_min = ga.p_min, p_max = ga.p_max;
  if (ga.skip && *ga.skip) return;
  __shared__ double red[kGcvThreads / 64][kGrid][4];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  constexpr int WAVES = kGcvThreads / 64;
  if (vb.bx == 0 && tid == 0) gst[3].nb[0] = vb.gx;
  const int chunk = (N + vb.gx - 1) / vb.gx;
  const int k0 = vb.bx * chunk, k1 = min(N, k0 + chunk);
  const double s = pow(10.0, grid_p(lane, p_min, p_max));
  double acc[4] = {0, 0, 0, 0};
  for (int k = k0 + wave; k < k1; k += 4 * WAVES) {
    double S4[4], L4[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int kk = k + e * WAVES;
      const bool in = kk < k1;
      const double y0 = in ? (double)Y[kk] : 0.0, y1 = in ? (double)Y[N + kk] : 0.0;
      S4[e] = y0 * y0 + y1 * y1;
      L4[e] = in ? lam[kk] : 0.0;
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      if (k + e * WAVES < k1) {
        double t[6];
        gcv_terms(S4[e], L4[e], s, t);
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[q] += t[q];
      }
    }
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) red[wave][lane][q] = acc[q];
  __syncthreads();
  const int p = tid & 63, q = tid >> 6;
  double v = 0.0;
#pragma unroll
  for (int w = 0; w < WAVES; ++w) v += red[w][p][q];
  part_base[kGcvPart0 + (size_t)vb.bx * (kGrid * 4) + p * 4 + q] = v;
}

// Second stage of the forward DCT fused with the grid evaluation: workgroup
// (bx, by) computes the 16 x 16 tile T1_f = Cm T0_f of both fields, then
// scores the 64 grid p-values over the tile's elements (lane j of each wave
// takes p-value j, the four waves every fourth element) -> partials
// part0[tile][p][sum], as gcv_grid_kernel would over the finished spectrum.
template <typename R>
struct Dct2GcvArgs {
  const R* cm;
  const R* t0;
  R* t1;
  int m, n;
  const double* lam;
  GcvState* gst;
  double* part_base;
  double p_min, p_max;
};
template <typename R>
struct Dct2GcvBody {
  using Args = Dct2GcvArgs<R>;
  static __device__ void run(const Args& a, VB vb);
};
template <typename R>
__global__ __launch_bounds__(256) void dct2_gcv_kernel(Dct2GcvArgs<R> a) {
  Dct2GcvBody<R>::run(a, vb_native());
}
template <typename R>
__device__ void Dct2GcvBody<R>::run(const Args& da, VB vb) {
  const R* cm = da.cm;
  const R* t0 = da.t0;
  R* t1 = da.t1;
  const int m = da.m, n = da.n;
  const double* lam = da.lam;
  GcvState* gst = da.gst;
  double* part_base = da.part_base;
  const double p_min = da.p_min, p_max = da.p_max;
  constexpr int KC = 64;
  __shared__ R As[16][KC + 1];
  __shared__ R Bs[2][KC][17];
  __shared__ double sS[256], sL[256];
  __shared__ double red[4][kGrid][4];
  const int N = m * n;
  const int tid = threadIdx.x, ty = tid / 16, tx = tid % 16;
  const int r0 = vb.by * 16, c0 = vb.bx * 16;
  const int tile = vb.by * vb.gx + vb.bx;
  if (tile == 0 && tid == 0) gst[3].nb[0] = vb.gx * vb.gy;
  R acc0 = R(0), acc1 = R(0);
  for (int k0 = 0; k0 < m; k0 += KC) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int e = tid + 256 * i;
      {
        const int rr = e / KC, ka = e % KC;  // A = Cm rows r0.., columns k0..
        const int r = r0 + rr, k = k0 + ka;
        As[rr][ka] = (r < m && k < m) ? cm[(size_t)r * m + k] : R(0);
      }
      {
        const int kb = e / 16, cc = e % 16;  // B = T0 rows k0.., columns c0..
        const int k = k0 + kb, c = c0 + cc;
        const bool in = k < m && c < n;
        Bs[0][kb][cc] = in ? t0[(size_t)k * n + c] : R(0);
        Bs[1][kb][cc] = in ? t0[(size_t)N + (size_t)k * n + c] : R(0);
      }
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < KC; ++kk) {
      const R a = As[ty][kk];
      acc0 = sfma(a, Bs[0][kk][tx], acc0);
      acc1 = sfma(a, Bs[1][kk][tx], acc1);
    }
    __syncthreads();
  }
  const int r = r0 + ty, c = c0 + tx;
  const bool in = r < m && c < n;
  if (in) {
    t1[(size_t)r * n + c] = acc0;
    t1[(size_t)N + (size_t)r * n + c] = acc1;
  }
  // out-of-range elements are marked by lam < 0 and skipped
  sS[tid] = (double)acc0 * (double)acc0 + (double)acc1 * (double)acc1;
  sL[tid] = in ? lam[(size_t)r * n + c] : -1.0;
  __syncthreads();
  const int lane = tid & 63, wave = tid >> 6;
  const double s = pow(10.0, grid_p(lane, p_min, p_max));
  double a4[4] = {0, 0, 0, 0};
  // 64 elements per wave in groups of eight independent evaluations (their
  // divisions overlap), each group summed as a tree before it is added
#pragma unroll 1
  for (int e0 = wave; e0 < 256; e0 += 32) {
    double g[8][4];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const double L = sL[e0 + 4 * i];
      double t[6];
      gcv_terms(sS[e0 + 4 * i], L >= 0.0 ? L : 0.0, s, t);
#pragma unroll
      for (int q = 0; q < 4; ++q) g[i][q] = L >= 0.0 ? t[q] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q)
      a4[q] += ((g[0][q] + g[1][q]) + (g[2][q] + g[3][q])) + ((g[4][q] + g[5][q]) + (g[6][q] + g[7][q]));
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) red[wave][lane][q] = a4[q];
  __syncthreads();
  const int p = tid & 63, q = tid >> 6;
  part_base[kGcvPart0 + (size_t)tile * (kGrid * 4) + p * 4 + q] = red[0][p][q] + red[1][p][q] + red[2][p][q] +
                                                                     red[3][p][q];
}

// Newton step STEP: decide p from the previous launch's partials, then score
// it over this workgroup's part of the spectrum.
template <int STEP, typename R>
struct NewtonBody {
  using Args = GcvArgs<R>;
  static __device__ void run(const Args& a, VB vb);
};
template <int STEP, typename R>
__global__ __launch_bounds__(kGcvThreads) void gcv_newton_kernel(GcvArgs<R> a) {
  NewtonBody<STEP, R>::run(a, vb_native());
}
template <int STEP, typename R>
__device__ void NewtonBody<STEP, R>::run(const Args& ga, VB vb) {
  const R* Y = ga.Y;
  const double* lam = ga.lam;
  const int N = ga.N;
  GcvState* gst = ga.gst;
  double* part_base = ga.part_base;
  const double p_min = ga.p_min, p_max = ga.p_max;
  if (ga.skip && *ga.skip) return;
  __shared__ GcvState ls;
  __shared__ double tot[kGrid][6];
  __shared__ double red[kGcvThreads / 64][6];
  gcv_decide_from<STEP - 1>(gst, part_base, N, p_min, p_max, &ls, tot);
  if (vb.bx == 0 && threadIdx.x == 0) {
    gcv_publish<STEP - 1>(gst, ls);
    gst[3].nb[STEP] = vb.gx;
  }
  if (ls.done) return;  // uniform: the minimum is fixed
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int chunk = (N + vb.gx - 1) / vb.gx;
  const int k0 = vb.bx * chunk, k1 = min(N, k0 + chunk);
  const double s = pow(10.0, ls.pv[0]);
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (int k = k0 + tid; k < k1; k += 4 * kGcvThreads) {
    double S4[4], L4[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int kk = k + e * kGcvThreads;
      const bool in = kk < k1;
      const double y0 = in ? (double)Y[kk] : 0.0, y1 = in ? (double)Y[N + kk] : 0.0;
      S4[e] = y0 * y0 + y1 * y1;
      L4[e] = in ? lam[kk] : 0.0;
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      if (k + e * kGcvThreads < k1) {
        double t[6];
        gcv_terms(S4[e], L4[e], s, t);
#pragma unroll
        for (int q = 0; q < 6; ++q) acc[q] += t[q];
      }
    }
  }
#pragma unroll
  for (int q = 0; q < 6; ++q) acc[q] = wave_sum(acc[q]);
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < 6; ++q) red[wave][q] = acc[q];
  }
  __syncthreads();
  if (tid < 6)
    part_base[(STEP == 1 ? kGcvPart1 : kGcvPart2) + (size_t)vb.bx * 6 + tid] =
        red[0][tid] + red[1][tid] + red[2][tid] + red[3][tid];
}

// The final decision alone (for consumers that read st->p).
struct GcvFinalBody {
  using Args = GcvArgs<float>;  // Y and lam unused
  static __device__ void run(const Args& a, VB) {
    if (a.skip && *a.skip) return;
    __shared__ GcvState ls;
    __shared__ double tot[kGrid][6];
    gcv_decide_from<2>(a.gst, a.part_base, a.N, a.p_min, a.p_max, &ls, tot);
    if (threadIdx.x == 0) gcv_publish<2>(a.gst, ls);
  }
};
__global__ __launch_bounds__(64) void gcv_final_kernel(GcvArgs<float> a) { GcvFinalBody::run(a, vb_native()); }

// The GCV search on the spectrum Y (two fields, 2N floats).  With
// final_launch the decided p is in st->p when the launches end; without it
// the consumer decides it (gcv_decide_from<2>).
// Recording of a launch (FP32 only: batches are FP32).
template <typename R, typename A>
bool recorded(LaunchKind kind, const A& a, dim3 grid, int block) {
  if constexpr (sizeof(R) != 4) return false;
  else return record_launch(kind, a, grid, block);
}

template <typename R>
void enqueue_gcv_newton(const R* Y, const double* lam, int N, GcvState* st, double* part, double p_min,
                        double p_max, const int32_t* skip, hipStream_t stream) {
  const dim3 et(kGcvThreads);
  const int nb = std::max(1, std::min(kGcvNewtonBlocks, (N + 4 * kGcvThreads - 1) / (4 * kGcvThreads)));
  const GcvArgs<R> a{Y, lam, N, st, part, p_min, p_max, skip};
  if (!recorded<R>(LaunchKind::kNewton1, a, dim3(nb), kGcvThreads))
    hipLaunchKernelGGL((gcv_newton_kernel<1, R>), dim3(nb), et, 0, stream, a);
  if (!recorded<R>(LaunchKind::kNewton2, a, dim3(nb), kGcvThreads))
    hipLaunchKernelGGL((gcv_newton_kernel<2, R>), dim3(nb), et, 0, stream, a);
}

template <typename R>
void enqueue_gcv_search(const R* Y, const double* lam, int N, GcvState* st, double* part, double p_min,
                        double p_max, const int32_t* skip, hipStream_t stream, bool final_launch = true) {
  const dim3 et(kGcvThreads);
  const int gb = std::max(1, std::min(kGcvBlocks, (N + 127) / 128));
  const GcvArgs<R> a{Y, lam, N, st, part, p_min, p_max, skip};
  if (!recorded<R>(LaunchKind::kGcvGrid, a, dim3(gb), kGcvThreads))
    hipLaunchKernelGGL((gcv_grid_kernel<R>), dim3(gb), et, 0, stream, a);
  enqueue_gcv_newton(Y, lam, N, st, part, p_min, p_max, skip, stream);
  if (final_launch) {
    const GcvArgs<float> f{nullptr, nullptr, N, st, part, p_min, p_max, skip};
    if (!recorded<R>(LaunchKind::kGcvFinal, f, dim3(1), 64))
      hipLaunchKernelGGL(gcv_final_kernel, dim3(1), dim3(64), 0, stream, f);
  }
}

// ---------------------------------------------------------------------------
// Small batched GEMM for the 2-D DCTs: for f in {0, 1}
//   C_f(r, c) = sum_k A_f(r, k) B_f(k, c),  A(r,k) = ta ? A[k*lda + r] : A[r*lda + k]
//                                           B(k,c) = tb ? B[c*ldb + k] : B[k*ldb + c]
// 16 x 16 output tile per 256-thread workgroup (one output per thread), K
// staged through LDS in chunks of 64; FMA accumulation in k order.
// ---------------------------------------------------------------------------
template <typename R>
struct GemmArgs {
  const R* a[2];
  const R* b[2];
  R* c[2];
  int lda, ldb, ldc;
  int ta, tb;
  int M, N, K;
  // optional: B(k, c) scaled by Gamma = 1 / (1 + 10^p lam[k * ldb + c])
  // (untransposed B), p fixed or decided here from the GCV search's last
  // partials (gcv_decide_from<2>); p written to p_out
  const double* lam;
  GcvState* gst;
  double p_fixed;
  double* p_out;
  const double* gpart;
  int gN;
  double gpmin, gpmax;
};

template <typename R>
struct GemmBody {
  using Args = GemmArgs<R>;
  static __device__ void run(const Args& g, VB vb);
};
template <typename R>
__global__ __launch_bounds__(256) void gemm_small_kernel(GemmArgs<R> g) {
  GemmBody<R>::run(g, vb_native());
}
template <typename R>
__device__ void GemmBody<R>::run(const Args& g, VB vb) {
  constexpr int KC = 64;  // K per LDS stage: four loads of A and four of B in flight per thread
  __shared__ R As[16][KC + 1];
  __shared__ R Bs[KC][17];
  const int f = vb.bz;
  const R* A = g.a[f];
  const R* B = g.b[f];
  const int r0 = vb.by * 16, c0 = vb.bx * 16;
  const int ty = threadIdx.x / 16, tx = threadIdx.x % 16;
  double sgam = 0.0;
  if (g.lam) {
    double p = g.p_fixed;
    const bool first = vb.bx == 0 && vb.by == 0 && vb.bz == 0 && threadIdx.x == 0;
    if (g.gst) {
      __shared__ GcvState ls;
      __shared__ double tot[kGrid][6];
      gcv_decide_from<2>(g.gst, g.gpart, g.gN, g.gpmin, g.gpmax, &ls, tot);
      p = ls.p;
      if (first) gcv_publish<2>(g.gst, ls);
    }
    sgam = pow(10.0, p);
    if (g.p_out && first) *g.p_out = p;
  }
  R acc = R(0);
  for (int k0 = 0; k0 < g.K; k0 += KC) {
    R va[4], vb[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      // consecutive threads run along the contiguous direction of each operand
      const int e = threadIdx.x + 256 * i;
      const int rr = g.ta ? e % 16 : e / KC, ka = g.ta ? e / 16 : e % KC;
      const int r = r0 + rr, k = k0 + ka;
      va[i] = (r < g.M && k < g.K) ? (g.ta ? A[(size_t)k * g.lda + r] : A[(size_t)r * g.lda + k]) : R(0);
      const int kb = g.tb ? e % KC : e / 16, cc = g.tb ? e / KC : e % 16;
      const int k2 = k0 + kb, c = c0 + cc;
      vb[i] = (k2 < g.K && c < g.N) ? (g.tb ? B[(size_t)c * g.ldb + k2] : B[(size_t)k2 * g.ldb + c]) : R(0);
      if (g.lam && k2 < g.K && c < g.N) {  // z_dct = Gamma f_dct (gpu_smoothn.py:196-199)
        const double G = 1.0 / (1.0 + sgam * g.lam[(size_t)k2 * g.ldb + c]);
        vb[i] = (R)(G * (double)vb[i]);
      }
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int e = threadIdx.x + 256 * i;
      As[g.ta ? e % 16 : e / KC][g.ta ? e / 16 : e % KC] = va[i];
      Bs[g.tb ? e % KC : e / 16][g.tb ? e / KC : e % 16] = vb[i];
    }
    __syncthreads();
    // a whole chunk, unrolled: the staged operands are zero beyond K, so the
    // extra terms add exact zeros and the LDS loads batch ahead of the FMAs
#pragma unroll
    for (int kk = 0; kk < KC; ++kk) acc = sfma(As[ty][kk], Bs[kk][tx], acc);
    __syncthreads();
  }
  const int r = r0 + ty, c = c0 + tx;
  if (r < g.M && c < g.N) g.c[f][(size_t)r * g.ldc + c] = acc;
}

template <typename R>
hipError_t gemm2(const R* a0, const R* a1, int lda, int ta, const R* b0, const R* b1, int ldb, int tb, R* c0, R* c1,
                 int ldc, int M, int N, int K, hipStream_t s, const double* lam = nullptr,
                 GcvState* gst = nullptr, double p_fixed = 0.0, double* p_out = nullptr,
                 const double* gpart = nullptr, int gN = 0, double gpmin = 0.0, double gpmax = 0.0) {
  GemmArgs<R> g{{a0, a1}, {b0, b1}, {c0, c1}, lda, ldb, ldc, ta, tb, M, N, K, lam, gst, p_fixed, p_out,
                gpart, gN, gpmin, gpmax};
  dim3 grid((N + 15) / 16, (M + 15) / 16, 2);
  if (recorded<R>(LaunchKind::kGemm, g, grid, 256)) return hipSuccess;
  hipLaunchKernelGGL(gemm_small_kernel<R>, grid, dim3(256), 0, s, g);
  return hipGetLastError();
}

}  // namespace

constexpr size_t kGcvPartOffset = (kGcvSlots * sizeof(GcvState) + 255) & ~size_t(255);
size_t smoothn_state_bytes() { return kGcvPartOffset + kGcvPartTotal * sizeof(double); }

hipError_t launch_smoothn_batch(LaunchKind kind, const void* args, const int4* grids, int n, int total_blocks,
                                hipStream_t s) {
  switch (kind) {
    case LaunchKind::kGemm: return launch_batched<GemmBody<float>, 256>(args, grids, n, total_blocks, s);
    case LaunchKind::kDct2Gcv: return launch_batched<Dct2GcvBody<float>, 256>(args, grids, n, total_blocks, s);
    case LaunchKind::kGcvGrid: return launch_batched<GcvGridBody<float>, kGcvThreads>(args, grids, n, total_blocks, s);
    case LaunchKind::kNewton1: return launch_batched<NewtonBody<1, float>, kGcvThreads>(args, grids, n, total_blocks, s);
    case LaunchKind::kNewton2: return launch_batched<NewtonBody<2, float>, kGcvThreads>(args, grids, n, total_blocks, s);
    case LaunchKind::kGcvFinal: return launch_batched<GcvFinalBody, 64>(args, grids, n, total_blocks, s);
    default: return hipErrorInvalidValue;
  }
}

template <typename R>
hipError_t launch_smoothn(R* u, R* v, int m, int n, double p_min, double p_max, double fixed_p, int auto_s,
                          const SmoothnWorkT<R>& w, hipStream_t stream) {
  const int N = m * n;
  if (N < 2) return hipSuccess;  // gpu_smoothn.py:137-140
  hipError_t e;
  // forward orthonormal DCT-II: T0 = X Cn^T ; T1 = Cm T0
  e = gemm2(u, v, n, 0, w.cos_n, w.cos_n, n, 1, w.t0, w.t0 + N, n, m, n, n, stream);
  if (e != hipSuccess) return e;
  // tiles of the fused second stage: one partial block each, and the
  // search state holds kGcvBlocks of them
  const int tiles = ((n + 15) / 16) * ((m + 15) / 16);
  if (auto_s && tiles <= kGcvBlocks) {
    // T1 = Cm T0 with the GCV grid evaluation over each tile, then the
    // Newton steps; the inverse DCT below makes the final decision
    GcvState* st = reinterpret_cast<GcvState*>(w.state);
    double* part = reinterpret_cast<double*>(reinterpret_cast<char*>(w.state) + kGcvPartOffset);
    const Dct2GcvArgs<R> da{w.cos_m, (const R*)w.t0, w.t1, m, n, w.lambda, st, part, p_min, p_max};
    const dim3 grid((n + 15) / 16, (m + 15) / 16);
    if (!recorded<R>(LaunchKind::kDct2Gcv, da, grid, 256))
      hipLaunchKernelGGL((dct2_gcv_kernel<R>), grid, dim3(256), 0, stream, da);
    enqueue_gcv_newton(w.t1, w.lambda, N, st, part, p_min, p_max, nullptr, stream);
  } else if (auto_s) {
    // larger fields: the plain second stage, then the grid evaluation over
    // at most kGcvBlocks workgroups
    e = gemm2(w.cos_m, w.cos_m, m, 0, w.t0, w.t0 + N, n, 0, w.t1, w.t1 + N, n, m, n, m, stream);
    if (e != hipSuccess) return e;
    GcvState* st = reinterpret_cast<GcvState*>(w.state);
    double* part = reinterpret_cast<double*>(reinterpret_cast<char*>(w.state) + kGcvPartOffset);
    enqueue_gcv_search(w.t1, w.lambda, N, st, part, p_min, p_max, nullptr, stream, false);
  } else {
    e = gemm2(w.cos_m, w.cos_m, m, 0, w.t0, w.t0 + N, n, 0, w.t1, w.t1 + N, n, m, n, m, stream);
    if (e != hipSuccess) return e;
  }
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  // inverse: T0 = Cm^T (Gamma T1) ; Z = T0 Cn  (the Gamma scaling is applied
  // while T1 is read)
  e = gemm2(w.cos_m, w.cos_m, m, 1, w.t1, w.t1 + N, n, 0, w.t0, w.t0 + N, n, m, n, m, stream, w.lambda,
            auto_s ? reinterpret_cast<GcvState*>(w.state) : nullptr, fixed_p, w.p_out,
            reinterpret_cast<const double*>(reinterpret_cast<const char*>(w.state) + kGcvPartOffset), N, p_min,
            p_max);
  if (e != hipSuccess) return e;
  return gemm2(w.t0, w.t0 + N, n, 0, w.cos_n, w.cos_n, n, 0, u, v, n, m, n, n, stream);
}


// ===========================================================================
// Weighted smoothing for masked or missing vectors (SmoothnGPU with a mask,
// gpu_smoothn.py:123-297): missing values take the value of their nearest
// finite neighbour, weights are 1 at finite unmasked vectors and 0
// elsewhere, z starts from a coarse DCT smoothing, and the relaxed update
//   z <- r IDCT2(Gamma DCT2(W (f - z) + z)) + (1 - r) z,   r = 1.75
// runs until ||dz|| / (||z|| + 1e-6) <= 1e-3 or 100 iterations, with the
// GCV search re-run at the iterations k where k + 1 is a power of two.
//
// Every iteration is three launches: a row transform with the W (f - z) + z
// prologue, a column-strip kernel doing the column DCT, the Gamma scaling
// and the column inverse in LDS, and a row inverse with the relaxation
// epilogue whose last block to finish reduces the change norms and raises
// the converged flag.  All later launches read that flag and return at
// once, so the host enqueues the full iteration budget with no
// synchronisation (and the sequence can sit in a graph).
// ===========================================================================
namespace {

constexpr int kMaxIters = 100;   // MAX_SMOOTHING_ITERS, gpu_smoothn.py:23
constexpr double kTolZ = 1e-3;   // TOL, gpu_smoothn.py:24
constexpr float kRelax = 1.75f;  // 1 + 0.75 * is_weighted, gpu_smoothn.py:162
// rows per block of the row transforms and columns per block of the column
// transforms (LDS: kSH x kMaxDim and 2 x kMaxDim x kSW values)
template <typename R> constexpr int kSH = sizeof(R) == 8 ? 4 : 8;
template <typename R> constexpr int kSW = sizeof(R) == 8 ? 2 : 8;
constexpr int kMaxDim = 1024;

struct WState {
  int32_t converged;   // the relative change of z reached the tolerance
  int32_t result;      // ping-pong slot of the last update
  int32_t iterations;
  uint32_t arrivals;   // blocks of the current update launch that finished
  double tol;
};

// Nearest finite row of each element within its column, ties to the upper
// one; -1 for a column with no finite element.  One 256-thread block per
// column: the last finite row at or above (max-scan) and the first at or
// below (min-scan) by Hillis-Steele scans in LDS (m <= kMaxDim).
template <typename R>
__global__ __launch_bounds__(256) void nn_column_kernel(const R* u, const R* v, int m, int n, int32_t* nr) {
  __shared__ int up[2][1024];
  __shared__ int dn[2][1024];
  const int j = blockIdx.x;
  for (int i = threadIdx.x; i < m; i += blockDim.x) {
    const size_t k = (size_t)i * n + j;
    const bool fin = isfinite(u[k]) && isfinite(v[k]);
    up[0][i] = fin ? i : -1;
    dn[0][i] = fin ? i : INT_MAX;
  }
  __syncthreads();
  int cur = 0;
  for (int off = 1; off < m; off <<= 1) {
    for (int i = threadIdx.x; i < m; i += blockDim.x) {
      const int a = up[cur][i], b = dn[cur][i];
      up[cur ^ 1][i] = i >= off ? max(a, up[cur][i - off]) : a;
      dn[cur ^ 1][i] = i + off < m ? min(b, dn[cur][i + off]) : b;
    }
    __syncthreads();
    cur ^= 1;
  }
  for (int i = threadIdx.x; i < m; i += blockDim.x) {
    const int a = up[cur][i], b = dn[cur][i];
    nr[(size_t)i * n + j] = (b != INT_MAX && (a < 0 || b - i < i - a)) ? b : a;
  }
}

// Weights, and the data with every non-finite vector replaced by the
// vector at the nearest finite position (exact Euclidean distance, unit
// spacing; equal distances resolved to the smaller column, then the smaller
// row -- the choice of scipy's exact feature transform used at
// gpu_smoothn.py:244-251).  The column scan stops once |dj|^2 exceeds the
// best squared distance.  Also resets the iteration state.
template <typename R>
__global__ __launch_bounds__(256) void nn_fill_kernel(const R* u, const R* v, const uint8_t* mask, int m, int n,
                                                      const int32_t* nr, R* f, R* w, WState* st) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k == 0) {
    st->converged = 0;
    st->result = 0;
    st->iterations = 0;
    st->arrivals = 0;
    st->tol = 0.0;
  }
  const int N = m * n;
  if (k >= N) return;
  const int i = k / n, j = k % n;
  const R a = u[k], b = v[k];
  const bool fin = isfinite(a) && isfinite(b);
  w[k] = (fin && !(mask && mask[k])) ? R(1) : R(0);
  if (fin) {
    f[k] = a;
    f[N + k] = b;
    return;
  }
  long long best = LLONG_MAX;
  int bi = -1, bj = -1;
  for (int d = 0; d < n; ++d) {
    if ((long long)d * d > best) break;
    for (int side = 0; side < 2; ++side) {
      if (d == 0 && side) break;
      const int jj = side ? j + d : j - d;
      if (jj < 0 || jj >= n) continue;
      const int r = nr[(size_t)i * n + jj];
      if (r < 0) continue;
      const long long d2 = (long long)d * d + (long long)(r - i) * (r - i);
      if (d2 < best || (d2 == best && (jj < bj || (jj == bj && r < bi)))) {
        best = d2;
        bi = r;
        bj = jj;
      }
    }
  }
  // no finite vector at all: the weights are all zero and the value is never
  // used unweighted; 0 keeps the transforms finite
  f[k] = bi < 0 ? R(0) : u[(size_t)bi * n + bj];
  f[N + k] = bi < 0 ? R(0) : v[(size_t)bi * n + bj];
}

__device__ __forceinline__ double block_sum_d(double x, double* red) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wave] = x;
  __syncthreads();
  double t = 0.0;
  for (int q = 0; q < (int)(blockDim.x >> 6); ++q) t += red[q];
  return t;
}

// Row transforms out[r][c] = sum_j X[r][j] M[j][c] over kSH rows per block
// (blockIdx.y = field, blockIdx.z = p-value of a batched GCV evaluation).
template <typename R>
struct RowArgs {
  const R* in;          // pro 0 input (2N per batch entry)
  const R* f;           // pro 1 / epi 2: filled data (2N)
  const R* w;           // weights (N)
  const R* z;           // current z (2N): pro 1, epi 1 relaxation
  const R* M;           // n x n
  R* out;               // 2N (epi 0, 1)
  int m, n;
  int pro;              // 0: in; 1: w (f - z) + z
  int epi;              // 0: store; 1: relaxed update (+ change norms); 2: GCV residual partials
  R relax;
  int track;            // epi 1: change norms, last block decides convergence
  int slot;             // epi 1: ping-pong slot written
  int iteration;
  const int32_t* skip;
  WState* st;
  double* part;         // epi 1: 2 per block; epi 2: (p, field, block) x {rss, trace}
  const GcvState* gst;  // epi 2: p-values
  const double* lambda; // epi 2
};

template <typename R>
__global__ __launch_bounds__(256) void rows_kernel(RowArgs<R> a) {
  constexpr int SH = kSH<R>;
  if (a.skip && *a.skip) return;
  __shared__ R X[SH][kMaxDim];
  __shared__ double red[4];
  __shared__ int is_last;
  const int field = blockIdx.y, pz = blockIdx.z;
  const int m = a.m, n = a.n;
  const size_t N = (size_t)m * n;
  const int r0 = blockIdx.x * SH;
  const R* in = a.in ? a.in + (size_t)pz * 2 * N + (size_t)field * N : nullptr;
  for (int e = threadIdx.x; e < SH * n; e += blockDim.x) {
    const int rr = e / n, jc = e - rr * n, r = r0 + rr;
    R x = R(0);
    if (r < m) {
      const size_t k = (size_t)r * n + jc;
      if (a.pro == 0) {
        x = in[k];
      } else {
        const R zz = a.z[field * N + k];
        x = a.w[k] * (a.f[field * N + k] - zz) + zz;
      }
    }
    X[rr][jc] = x;
  }
  __syncthreads();
  double s0 = 0.0, s1 = 0.0;
  double trace = 0.0;
  const double sp = a.epi == 2 ? pow(10.0, a.gst->pv[pz]) : 0.0;
  for (int c = threadIdx.x; c < n; c += blockDim.x) {
    R acc[SH];
#pragma unroll
    for (int rr = 0; rr < SH; ++rr) acc[rr] = R(0);
    for (int jj = 0; jj < n; ++jj) {
      const R mv = a.M[(size_t)jj * n + c];
#pragma unroll
      for (int rr = 0; rr < SH; ++rr) acc[rr] = sfma(X[rr][jj], mv, acc[rr]);
    }
#pragma unroll
    for (int rr = 0; rr < SH; ++rr) {
      const int r = r0 + rr;
      if (r >= m) break;
      const size_t k = (size_t)r * n + c;
      if (a.epi == 0) {
        a.out[field * N + k] = acc[rr];
      } else if (a.epi == 1) {
        R zn = acc[rr];
        if (a.relax != R(1)) {
          const R zo = a.z[field * N + k];
          zn = a.relax * acc[rr] + (R(1) - a.relax) * zo;
          const double dz = (double)zo - (double)zn;
          s0 += dz * dz;
        }
        s1 += (double)zn * (double)zn;
        a.out[field * N + k] = zn;
      } else {
        const double res = (double)a.f[field * N + k] - (double)acc[rr];
        s0 += (double)a.w[k] * res * res;
        if (field == 0) trace += 1.0 / (1.0 + sp * a.lambda[k]);
      }
    }
  }
  if (a.epi == 0 || (a.epi == 1 && !a.track)) return;
  const int nblk = gridDim.x;
  if (a.epi == 2) {
    s0 = block_sum_d(s0, red);
    trace = block_sum_d(trace, red);
    if (threadIdx.x == 0) {
      double* o = a.part + (((size_t)pz * 2 + field) * nblk + blockIdx.x) * 2;
      o[0] = s0;
      o[1] = trace;
    }
    return;
  }
  s0 = block_sum_d(s0, red);
  s1 = block_sum_d(s1, red);
  const int bid = field * nblk + blockIdx.x, nb = nblk * gridDim.y;
  if (threadIdx.x == 0) {
    a.part[2 * bid] = s0;
    a.part[2 * bid + 1] = s1;
    __threadfence();
    is_last = atomicAdd(&a.st->arrivals, 1u) == (unsigned)(nb - 1);
  }
  __syncthreads();
  if (!is_last) return;
  __threadfence();
  if (threadIdx.x == 0) {
    // fixed order: the decision does not depend on block completion order
    double d2 = 0.0, z2 = 0.0;
    const volatile double* pp = a.part;
    for (int q = 0; q < nb; ++q) {
      d2 += pp[2 * q];
      z2 += pp[2 * q + 1];
    }
    const double tol = sqrt(d2) / (sqrt(z2) + 1e-6);
    a.st->tol = tol;
    a.st->result = a.slot;
    a.st->iterations = a.iteration + 1;
    a.st->arrivals = 0;
    if (tol <= kTolZ) a.st->converged = 1;
  }
}

// Column transforms on strips of kSW columns (blockIdx.y = field,
// blockIdx.z = p-value of a batched GCV evaluation):
//   mode 0: out = Cm^T (G .* (Cm X));  mode 1: out = Cm X;  mode 2: out = Cm^T (G .* X)
// G = Gamma(p) = 1 / (1 + 10^p Lambda), or the coarse-start truncation
// (1 where row < n0 or column < n1).  In-place use is safe: a block reads
// its whole strip before writing it.
template <typename R>
struct ColArgs {
  const R* in;
  R* out;               // + z * 2N when batched
  const R* Mf;          // forward, transposed: Mf[r * m + i] = C[i][r]
  const R* Mi;          // inverse: C[k][i] at k * m + i
  const double* lambda;
  int m, n;
  int mode;
  int trunc;
  int n0, n1;
  double p_fixed;
  const GcvState* gst;  // p = gst->p (or gst->pv[z] when batched) when non-null
  int batched;
  const int32_t* skip;
};

template <typename R>
__global__ __launch_bounds__(256) void cols_kernel(ColArgs<R> a) {
  constexpr int SW = kSW<R>;
  if (a.skip && *a.skip) return;
  __shared__ R X[kMaxDim][SW];
  __shared__ R Y[kMaxDim][SW];
  const int field = blockIdx.y, pz = blockIdx.z;
  const int m = a.m, n = a.n;
  const size_t N = (size_t)m * n;
  const int c0 = blockIdx.x * SW;
  const int nc = min(SW, n - c0);
  const R* in = a.in + (size_t)field * N;
  R* out = a.out + (size_t)(a.batched ? pz : 0) * 2 * N + (size_t)field * N;
  for (int e = threadIdx.x; e < m * SW; e += blockDim.x) {
    const int r = e / SW, c = e - r * SW;
    X[r][c] = c < nc ? in[(size_t)r * n + c0 + c] : R(0);
  }
  __syncthreads();
  double s = 0.0;
  if (!a.trunc) {
    const double p = a.gst ? (a.batched ? a.gst->pv[pz] : a.gst->p) : a.p_fixed;
    s = pow(10.0, p);
  }
  auto gain = [&](int i, int c) -> R {
    if (a.trunc) return (i < a.n0 || c0 + c < a.n1) ? R(1) : R(0);
    return (R)(1.0 / (1.0 + s * a.lambda[(size_t)i * n + c0 + c]));
  };
  if (a.mode != 2) {
    for (int i = threadIdx.x; i < m; i += blockDim.x) {
      R acc[SW];
#pragma unroll
      for (int c = 0; c < SW; ++c) acc[c] = R(0);
      for (int r = 0; r < m; ++r) {
        const R cv = a.Mf[(size_t)r * m + i];
#pragma unroll
        for (int c = 0; c < SW; ++c) acc[c] = sfma(cv, X[r][c], acc[c]);
      }
      if (a.mode == 1) {
        for (int c = 0; c < nc; ++c) out[(size_t)i * n + c0 + c] = acc[c];
      } else {
#pragma unroll
        for (int c = 0; c < SW; ++c) Y[i][c] = c < nc ? acc[c] * gain(i, c) : R(0);
      }
    }
    if (a.mode == 1) return;
  } else {
    for (int e = threadIdx.x; e < m * SW; e += blockDim.x) {
      const int r = e / SW, c = e - r * SW;
      Y[r][c] = c < nc ? X[r][c] * gain(r, c) : R(0);
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < m; i += blockDim.x) {
    R acc[SW];
#pragma unroll
    for (int c = 0; c < SW; ++c) acc[c] = R(0);
    for (int k = 0; k < m; ++k) {
      const R cv = a.Mi[(size_t)k * m + i];
#pragma unroll
      for (int c = 0; c < SW; ++c) acc[c] = sfma(cv, Y[k][c], acc[c]);
    }
    for (int c = 0; c < nc; ++c) out[(size_t)i * n + c0 + c] = acc[c];
  }
}

// Batched-residual GCV for w_mean <= 0.9 (gpu_smoothn.py:363-368): the
// score of each pending p-value from the partials of rows_kernel (epi 2);
// phase 0 scores the 10-point span, phases 1 and 2 a 64-point grid over the
// bracket around the previous minimum, and phase 2 ends with the vertex of
// the parabola through the minimum and its neighbours.
template <int PHASE>
__global__ __launch_bounds__(64) void gcv_slow_decide_kernel(int N, GcvState* st, const double* part, int nblk,
                                                             const int32_t* skip) {
  if (skip && *skip) return;
  const int lane = threadIdx.x;
  const int npv = st->npv;
  double g = 1e300, pv = 0.0;
  if (lane < npv) {
    double rss = 0.0, tr = 0.0;
    for (int fld = 0; fld < 2; ++fld)
      for (int b = 0; b < nblk; ++b) {
        const double* o = part + (((size_t)lane * 2 + fld) * nblk + b) * 2;
        rss += o[0];
        if (fld == 0) tr += o[1];
      }
    const double d = 1.0 - tr / N;
    g = rss / N / (d * d) / 2.0;
    pv = st->pv[lane];
  }
  double gb = g;
  int ib = lane;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const double go = __shfl_xor(gb, off, 64);
    const int io = __shfl_xor(ib, off, 64);
    if (go < gb || (go == gb && io < ib)) {
      gb = go;
      ib = io;
    }
  }
  const int il = ib > 0 ? ib - 1 : 0, ir = ib < npv - 1 ? ib + 1 : npv - 1;
  const double pl = __shfl(pv, il, 64), pb = __shfl(pv, ib, 64), pr = __shfl(pv, ir, 64);
  const double gl = __shfl(g, il, 64), gr = __shfl(g, ir, 64);
  __syncthreads();
  if (PHASE < 2) {
    st->pv[lane] = pl + (pr - pl) * (double)lane / (kGrid - 1);
    if (lane == 0) st->npv = kGrid;
  } else if (lane == 0) {
    double p = pb;
    if (ib > 0 && ib < npv - 1) {
      const double a1 = (pb - pl) * (gb - gr), a2 = (pb - pr) * (gb - gl);
      const double den = a1 - a2;
      if (den != 0.0) {
        const double q = pb - 0.5 * ((pb - pl) * a1 - (pb - pr) * a2) / den;
        if (q >= pl && q <= pr) p = q;
      }
    }
    st->p = p;
  }
}

__global__ void gcv_slow_init_kernel(GcvState* st, double p_min, double p_max, const int32_t* skip) {
  if (skip && *skip) return;
  const int i = threadIdx.x;
  if (i < kSpan) st->pv[i] = (double)i * (1.0 / (kSpan - 1)) * (p_max - p_min) + p_min;
  if (i == 0) {
    st->npv = kSpan;
    st->p_min = p_min;
    st->p_max = p_max;
  }
}

// z at the converged (or last) slot, 0 at masked vectors (:224-225).
template <typename R>
__global__ __launch_bounds__(256) void masked_out_kernel(R* u, R* v, const uint8_t* mask, const R* z, int N,
                                                         const WState* st, const GcvState* gst, double p_fixed,
                                                         double* p_out, int32_t* iters_out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const R* zz = z + (size_t)st->result * 2 * N;
  if (k < N) {
    const bool mk = mask && mask[k];
    u[k] = mk ? R(0) : zz[k];
    v[k] = mk ? R(0) : zz[N + k];
  }
  if (k == 0) {
    if (p_out) *p_out = gst ? gst->p : p_fixed;
    if (iters_out) *iters_out = st->iterations;
  }
}

// GCV partials: the spectral search's (three launches) or the
// batched-residual search's (kGrid p-values x 2 fields x up to 256 row
// blocks x 2), whichever is larger
constexpr size_t kGcvPartDoubles =
    kGcvPartTotal > (size_t)kGrid * 2 * 256 * 2 ? kGcvPartTotal : (size_t)kGrid * 2 * 256 * 2;
constexpr size_t kWStateOffset =
    (kGcvPartOffset + kGcvPartDoubles * sizeof(double) + 255) & ~size_t(255);

}  // namespace

size_t smoothn_masked_state_bytes() { return kWStateOffset + 256; }

size_t smoothn_masked_part_doubles(int m) { return 2 * 2 * (size_t)((m + kSH<double> - 1) / kSH<double>); }

int smoothn_masked_max_dim() { return kMaxDim; }

template <typename R>
hipError_t launch_smoothn_masked(R* u, R* v, int m, int n, double p_min, double p_max, double fixed_p, int auto_s,
                                 const SmoothnMaskedWorkT<R>& w, hipStream_t stream) {
  if (g_launch_record) return hipErrorNotSupported;  // masked smoothing is not recordable
  constexpr int SH = kSH<R>, SW = kSW<R>;
  const int N = m * n;
  if (N < 2) return hipSuccess;  // gpu_smoothn.py:137-140
  if (m > kMaxDim || n > kMaxDim) return hipErrorInvalidValue;
  char* base = reinterpret_cast<char*>(w.state);
  GcvState* gst = reinterpret_cast<GcvState*>(base);
  double* gpart = reinterpret_cast<double*>(base + kGcvPartOffset);
  WState* ws = reinterpret_cast<WState*>(base + kWStateOffset);
  const int32_t* skip = &ws->converged;
  const size_t N2 = 2 * (size_t)N;
  R* zs[2] = {w.z, w.z + N2};
  const dim3 rgrid((m + SH - 1) / SH, 2), cgrid((n + SW - 1) / SW, 2), tb(256);

  hipLaunchKernelGGL(nn_column_kernel<R>, dim3(n), dim3(256), 0, stream, u, v, m, n, w.nr);
  hipLaunchKernelGGL(nn_fill_kernel<R>, dim3((N + 255) / 256), dim3(256), 0, stream, u, v, w.mask, m, n, w.nr, w.f,
                     w.w, ws);

  RowArgs<R> ra{};
  ra.m = m;
  ra.n = n;
  ra.st = ws;
  ra.part = w.part;
  ColArgs<R> ca{};
  ca.m = m;
  ca.n = n;
  ca.Mf = w.cos_mt;
  ca.Mi = w.cos_m;
  ca.lambda = w.lambda;
  ca.p_fixed = fixed_p;

  // coarse start (gpu_smoothn.py:283-297): keep the coefficients with row <
  // ceil(m/10)+1 or column < ceil(n/10)+1
  RowArgs<R> r0 = ra;
  r0.in = w.f;
  r0.M = w.cos_nt;
  r0.out = w.t0;
  hipLaunchKernelGGL(rows_kernel<R>, rgrid, tb, 0, stream, r0);
  ColArgs<R> c0 = ca;
  c0.in = w.t0;
  c0.out = w.t1;
  c0.mode = 0;
  c0.trunc = 1;
  c0.n0 = (m + 9) / 10 + 1;
  c0.n1 = (n + 9) / 10 + 1;
  hipLaunchKernelGGL(cols_kernel<R>, cgrid, tb, 0, stream, c0);
  RowArgs<R> r1 = ra;
  r1.in = w.t1;
  r1.M = w.cos_n;
  r1.out = zs[0];
  r1.epi = 1;
  r1.relax = R(1);
  hipLaunchKernelGGL(rows_kernel<R>, rgrid, tb, 0, stream, r1);

  for (int k = 0; k < kMaxIters; ++k) {
    const R* zc = zs[k & 1];
    RowArgs<R> fa = ra;
    fa.f = w.f;
    fa.w = w.w;
    fa.z = zc;
    fa.pro = 1;
    fa.M = w.cos_nt;
    fa.out = w.t0;
    fa.skip = skip;
    hipLaunchKernelGGL(rows_kernel<R>, rgrid, tb, 0, stream, fa);
    const bool gcv = auto_s && ((k + 1) & k) == 0;
    ColArgs<R> cc = ca;
    cc.skip = skip;
    cc.gst = auto_s ? gst : nullptr;
    if (gcv) {
      ColArgs<R> cs = cc;
      cs.in = w.t0;
      cs.out = w.t1;
      cs.mode = 1;
      hipLaunchKernelGGL(cols_kernel<R>, cgrid, tb, 0, stream, cs);
      if (!w.slow_gcv) {
        enqueue_gcv_search(w.t1, w.lambda, N, gst, gpart, p_min, p_max, skip, stream);
      } else {
        hipLaunchKernelGGL(gcv_slow_init_kernel, dim3(1), dim3(64), 0, stream, gst, p_min, p_max, skip);
        for (int ph = 0; ph < 3; ++ph) {
          const int P = ph == 0 ? kSpan : kGrid;
          ColArgs<R> cb = cc;
          cb.in = w.t1;
          cb.out = w.qbatch;
          cb.mode = 2;
          cb.batched = 1;
          hipLaunchKernelGGL(cols_kernel<R>, dim3(cgrid.x, 2, P), tb, 0, stream, cb);
          RowArgs<R> rb = ra;
          rb.in = w.qbatch;
          rb.f = w.f;
          rb.w = w.w;
          rb.M = w.cos_n;
          rb.epi = 2;
          rb.gst = gst;
          rb.lambda = w.lambda;
          rb.part = gpart;
          rb.skip = skip;
          hipLaunchKernelGGL(rows_kernel<R>, dim3(rgrid.x, 2, P), tb, 0, stream, rb);
          if (ph == 0)
            hipLaunchKernelGGL(gcv_slow_decide_kernel<0>, dim3(1), dim3(64), 0, stream, N, gst, gpart, (int)rgrid.x,
                               skip);
          else if (ph == 1)
            hipLaunchKernelGGL(gcv_slow_decide_kernel<1>, dim3(1), dim3(64), 0, stream, N, gst, gpart, (int)rgrid.x,
                               skip);
          else
            hipLaunchKernelGGL(gcv_slow_decide_kernel<2>, dim3(1), dim3(64), 0, stream, N, gst, gpart, (int)rgrid.x,
                               skip);
        }
      }
      ColArgs<R> ci = cc;
      ci.in = w.t1;
      ci.out = w.t1;
      ci.mode = 2;
      hipLaunchKernelGGL(cols_kernel<R>, cgrid, tb, 0, stream, ci);
    } else {
      cc.in = w.t0;
      cc.out = w.t1;
      cc.mode = 0;
      hipLaunchKernelGGL(cols_kernel<R>, cgrid, tb, 0, stream, cc);
    }
    RowArgs<R> ru = ra;
    ru.in = w.t1;
    ru.M = w.cos_n;
    ru.z = zc;
    ru.out = zs[(k + 1) & 1];
    ru.epi = 1;
    ru.relax = (R)kRelax;
    ru.track = 1;
    ru.slot = (k + 1) & 1;
    ru.iteration = k;
    ru.skip = skip;
    hipLaunchKernelGGL(rows_kernel<R>, rgrid, tb, 0, stream, ru);
  }
  hipLaunchKernelGGL(masked_out_kernel<R>, dim3((N + 255) / 256), dim3(256), 0, stream, u, v, w.mask, w.z, N,
                     (const WState*)ws, auto_s ? (const GcvState*)gst : nullptr, fixed_p, w.p_out, w.iters_out);
  return hipGetLastError();
}

template hipError_t launch_smoothn<float>(float*, float*, int, int, double, double, double, int,
                                         const SmoothnWorkT<float>&, hipStream_t);
template hipError_t launch_smoothn<double>(double*, double*, int, int, double, double, double, int,
                                          const SmoothnWorkT<double>&, hipStream_t);
template hipError_t launch_smoothn_masked<float>(float*, float*, int, int, double, double, double, int,
                                                const SmoothnMaskedWorkT<float>&, hipStream_t);
template hipError_t launch_smoothn_masked<double>(double*, double*, int, int, double, double, double, int,
                                                 const SmoothnMaskedWorkT<double>&, hipStream_t);

}  // namespace piv
