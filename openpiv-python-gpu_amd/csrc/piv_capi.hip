// C ABI and pass-loop executor (PIVGPU.__call__, gpu_process.py:564-647).
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <climits>
#include <mutex>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/openpiv_hip.h"
#include "piv_cnn.h"
#include "piv_kernels.h"
#include "piv_record.h"

using namespace piv;

namespace {
thread_local std::string g_err;
}  // namespace

namespace piv {
int capi_fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}
}  // namespace piv

namespace {

#define fail piv::capi_fail

#define HIP_TRY(expr)                                                                               \
  do {                                                                                              \
    piv::set_launch_detail("", hipSuccess);                                                          \
    hipError_t _e = (expr);                                                                         \
    if (_e != hipSuccess)                                                                           \
      return fail(PIV_E_HIP, "%s failed: %s [%s]", #expr, hipGetErrorString(_e), piv::launch_detail()); \
  } while (0)

bool is_pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }

// Kernel half-widths of a pass (0 reads as the reference's default 1).
int validation_size(const piv_pass_desc& p) { return p.validation_size > 0 ? p.validation_size : 1; }
int replacing_size(const piv_pass_desc& p) { return p.replacing_size > 0 ? p.replacing_size : 1; }

// Validation / replacement settings of a pass on an m x n field.
int check_tests(const piv_pass_desc& p, int m, int n) {
  if (p.s2n_method < PIV_PEAK2PEAK || p.s2n_method > PIV_PEAK2ENERGY) return fail(PIV_E_INVALID, "s2n_method");
  if (p.s2n_size < 1) return fail(PIV_E_INVALID, "s2n_size must be positive");
  if (p.replacing_method < PIV_SPRING || p.replacing_method > PIV_MEAN) return fail(PIV_E_INVALID, "replacing_method");
  if (p.num_replacing_iters < 0 || p.num_replacing_iters > PIV_MAX_REPLACING_ITERS)
    return fail(PIV_E_INVALID, "num_replacing_iters outside 0..%d", PIV_MAX_REPLACING_ITERS);
  if (p.validation_size < 0 || p.validation_size > kMaxNeighbourhoodSize)
    return fail(PIV_E_INVALID, "validation_size outside 1..%d", kMaxNeighbourhoodSize);
  if (p.replacing_size < 0 || p.replacing_size > kMaxNeighbourhoodSize)
    return fail(PIV_E_INVALID, "replacing_size outside 1..%d", kMaxNeighbourhoodSize);
  // gpu_validation.py:74, 273
  if (validation_size(p) > m || validation_size(p) > n || replacing_size(p) > m || replacing_size(p) > n)
    return fail(PIV_E_INVALID, "validation_size / replacing_size exceed the %d x %d field", m, n);
  return PIV_OK;
}

bool needs_s2n(const piv_pass_desc& p) { return p.validate && !std::isnan(p.s2n_tol); }

int check_pass(const piv_pass_desc& p, int ht, int wd) {
  if (p.search_size < 2 || p.spacing < 1) return fail(PIV_E_INVALID, "search_size/spacing out of range");
  if (!is_pow2(p.fft_size) || p.fft_size < 16 || p.fft_size > kMaxFft)
    return fail(PIV_E_INVALID, "fft_size %d unsupported (power of two in [16, %d])", p.fft_size, kMaxFft);
  if (p.search_size > p.fft_size || (p.fft_size - p.search_size) % 2)
    return fail(PIV_E_INVALID, "search_size %d incompatible with fft_size %d", p.search_size, p.fft_size);
  if (p.field_m < 1 || p.field_n < 1) return fail(PIV_E_INVALID, "empty field");
  if (p.x_offset < 0 || p.y_offset < 0 || (p.field_m - 1) * p.spacing + p.search_size + p.y_offset > ht ||
      (p.field_n - 1) * p.spacing + p.search_size + p.x_offset > wd)
    return fail(PIV_E_INVALID, "window grid exceeds the frame");
  if (p.subpixel < 0 || p.subpixel > 2) return fail(PIV_E_INVALID, "subpixel method");
  if (p.window_size != 0 && (p.window_size < 2 || p.window_size > p.search_size ||
                             (p.search_size - p.window_size) % 2))
    return fail(PIV_E_INVALID, "window_size %d incompatible with search_size %d", p.window_size, p.search_size);
  return check_tests(p, p.field_m, p.field_n);
}

size_t dtype_size(int d) { return d == PIV_U8 ? 1 : (d == PIV_U16 ? 2 : (d == PIV_F32 ? 4 : 8)); }

// Grid node coordinates of a pass (gpu_process.py:940-947): x = j*sp + ss/2 + x_off.
float grid_x0(const piv_pass_desc& p) { return (float)p.search_size / 2.0f + (float)p.x_offset; }
float grid_y0(const piv_pass_desc& p) { return (float)p.search_size / 2.0f + (float)p.y_offset; }

// smoothn p bounds for order 2 (gpu_smoothn.py:320-338).
void p_bounds(int m, int n, double* pmin, double* pmax) {
  const double h_min = 1e-3, h_max = 1 - h_min;
  const int rank = (m != 1) + (n != 1);
  const double e = 2.0 / rank;
  auto f = [&](double h) {
    const double he = pow(h, e);
    const double t = (1 + sqrt(1 + 8 * he)) / 4 / he;
    return log10((t * t - 1) / 16);
  };
  *pmin = f(h_max);
  *pmax = f(h_min);
}

// Orthonormal DCT-II matrix C[k][i] = s_k cos(pi (2i+1) k / (2n)).
// FP32 tables round the FP64 value; FP64 tables round an extended-precision one.
template <typename R>
std::vector<R> dct_matrix(int n) {
  std::vector<R> c((size_t)n * n);
  for (int k = 0; k < n; ++k) {
    if (sizeof(R) == 4) {
      const double s = k == 0 ? sqrt(1.0 / n) : sqrt(2.0 / n);
      for (int i = 0; i < n; ++i) c[(size_t)k * n + i] = (R)(s * cos(M_PI * (2 * i + 1) * k / (2.0 * n)));
    } else {
      const long double s = k == 0 ? sqrtl(1.0L / n) : sqrtl(2.0L / n);
      for (int i = 0; i < n; ++i)
        c[(size_t)k * n + i] = (R)(s * cosl(3.14159265358979323846264338327950288L * (2 * i + 1) * k / (2.0L * n)));
    }
  }
  return c;
}

// Eigenvalues of the difference operator (gpu_smoothn.py:340-351), squared
// for smoothing order 2 (:165).
std::vector<double> lambda_table(int m, int n) {
  std::vector<double> l((size_t)m * n);
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < n; ++j) {
      const double v = 2.0 * (2.0 - (cos(M_PI * i / m) + cos(M_PI * j / n)));
      l[(size_t)i * n + j] = v * v;
    }
  return l;
}

// DCT matrices in the plan's value type (void*: float or double).
struct SmoothTables {
  void* cos_m = nullptr;
  void* cos_n = nullptr;
  void* cos_mt = nullptr;  // transposes (weighted path)
  void* cos_nt = nullptr;
  double* lambda = nullptr;
};

template <typename R>
std::vector<R> transposed(const std::vector<R>& c, int n) {
  std::vector<R> t(c.size());
  for (int a = 0; a < n; ++a)
    for (int b = 0; b < n; ++b) t[(size_t)b * n + a] = c[(size_t)a * n + b];
  return t;
}

template <typename R>
int upload(void** dst, const std::vector<R>& v) {
  HIP_TRY(hipMalloc(dst, v.size() * sizeof(R)));
  HIP_TRY(hipMemcpy(*dst, v.data(), v.size() * sizeof(R), hipMemcpyHostToDevice));
  return PIV_OK;
}

template <typename R>
int make_tables_t(int m, int n, SmoothTables* t) {
  auto cm = dct_matrix<R>(m), cn = dct_matrix<R>(n);
  int r = upload(&t->cos_m, cm);
  if (!r) r = upload(&t->cos_n, cn);
  if (!r) r = upload(&t->cos_mt, transposed(cm, m));
  if (!r) r = upload(&t->cos_nt, transposed(cn, n));
  if (!r) r = upload((void**)&t->lambda, lambda_table(m, n));
  return r;
}
int make_tables(int m, int n, SmoothTables* t, bool fp64 = false) {
  return fp64 ? make_tables_t<double>(m, n, t) : make_tables_t<float>(m, n, t);
}
void free_tables(SmoothTables* t) {
  for (void* p : {(void*)t->cos_m, (void*)t->cos_n, (void*)t->cos_mt, (void*)t->cos_nt, (void*)t->lambda})
    if (p) (void)hipFree(p);
  *t = SmoothTables{};
}

// Weighted-smoothing work buffers sized for fields of up to maxN vectors
// and maxM rows (slow: also the batched GCV residual buffer).
struct MaskedSmoothBufs {
  int32_t* nr = nullptr;
  void* f = nullptr;       // value-type arrays (float or double)
  void* w = nullptr;
  void* z = nullptr;
  double* part = nullptr;
  void* state = nullptr;
  void* qbatch = nullptr;
};

int alloc_masked_smooth(int maxN, int maxM, bool slow, MaskedSmoothBufs* b, size_t esize = 4) {
  HIP_TRY(hipMalloc(&b->nr, (size_t)maxN * 4));
  HIP_TRY(hipMalloc(&b->f, 2ull * maxN * esize));
  HIP_TRY(hipMalloc(&b->w, (size_t)maxN * esize));
  HIP_TRY(hipMalloc(&b->z, 4ull * maxN * esize));
  HIP_TRY(hipMalloc(&b->part, smoothn_masked_part_doubles(maxM) * 8));
  HIP_TRY(hipMalloc(&b->state, smoothn_masked_state_bytes()));
  HIP_TRY(hipMemset(b->state, 0, smoothn_masked_state_bytes()));
  if (slow) HIP_TRY(hipMalloc(&b->qbatch, 64ull * 2 * maxN * esize));
  return PIV_OK;
}
void free_masked_smooth(MaskedSmoothBufs* b) {
  for (void* p : {(void*)b->nr, (void*)b->f, (void*)b->w, (void*)b->z, (void*)b->part, b->state, (void*)b->qbatch})
    if (p) (void)hipFree(p);
  *b = MaskedSmoothBufs{};
}

template <typename R>
SmoothnMaskedWorkT<R> masked_work(const SmoothTables& tab, const uint8_t* fmask, const MaskedSmoothBufs& b, void* t0,
                                  void* t1, double* p_out, int32_t* iters_out, int slow) {
  SmoothnMaskedWorkT<R> w{};
  w.cos_m = (const R*)tab.cos_m;
  w.cos_mt = (const R*)tab.cos_mt;
  w.cos_n = (const R*)tab.cos_n;
  w.cos_nt = (const R*)tab.cos_nt;
  w.lambda = tab.lambda;
  w.mask = fmask;
  w.nr = b.nr;
  w.f = (R*)b.f;
  w.w = (R*)b.w;
  w.z = (R*)b.z;
  w.t0 = (R*)t0;
  w.t1 = (R*)t1;
  w.part = b.part;
  w.state = b.state;
  w.qbatch = (R*)b.qbatch;
  w.p_out = p_out;
  w.iters_out = iters_out;
  w.slow_gcv = slow;
  return w;
}

// Field mask of a pass: the frame mask at each window centre truncated to
// integer pixels, mask[int(y), int(x)] (gpu_process.py:949-951).
std::vector<uint8_t> field_mask_of(const piv_pass_desc& p, const uint8_t* mask, int wd) {
  std::vector<uint8_t> fm((size_t)p.field_m * p.field_n);
  const double x0 = p.search_size / 2.0 + p.x_offset, y0 = p.search_size / 2.0 + p.y_offset;
  for (int i = 0; i < p.field_m; ++i) {
    const int y = (int)(float)(i * (double)p.spacing + y0);
    for (int j = 0; j < p.field_n; ++j) {
      const int x = (int)(float)(j * (double)p.spacing + x0);
      fm[(size_t)i * p.field_n + j] = mask[(size_t)y * wd + x] ? 1 : 0;
    }
  }
  return fm;
}

// exp(-2 pi i k / F) for k < F/2 as (re, im) pairs, rounded from extended
// precision.
std::vector<double> fft_twiddles(int F) {
  std::vector<double> t((size_t)F);
  for (int k = 0; k < F / 2; ++k) {
    const long double ang = -2.0L * 3.14159265358979323846264338327950288L * k / F;
    t[2 * k] = (double)cosl(ang);
    t[2 * k + 1] = (double)sinl(ang);
  }
  return t;
}

CorrArgs corr_args(const void* fa, const void* fb, int dtype, int ht, int wd, const piv_pass_desc& p) {
  CorrArgs a{};
  a.frame_a = fa;
  a.frame_b = fb;
  a.in_dtype = dtype;
  a.ht = ht;
  a.wd = wd;
  a.ss = p.search_size;
  a.fsize = p.fft_size;
  a.spacing = p.spacing;
  a.x_off = p.x_offset;
  a.y_off = p.y_offset;
  a.m = p.field_m;
  a.n = p.field_n;
  a.normalize = p.normalize;
  a.subpixel = p.subpixel;
  a.s2n_method = p.s2n_method;
  a.s2n_size = p.s2n_size;
  a.win_pad = p.window_size > 0 ? (p.search_size - p.window_size) / 2 : 0;
  return a;
}

template <typename R>
DeformArgsT<R> deform_args(const void* fa, const void* fb, int dtype, int ht, int wd, const piv_pass_desc& p,
                           const R* up, const R* vp, R* oa, R* ob) {
  DeformArgsT<R> d{};
  d.frame_a = fa;
  d.frame_b = fb;
  d.in_dtype = dtype;
  d.ht = ht;
  d.wd = wd;
  d.up = up;
  d.vp = vp;
  d.m = p.field_m;
  d.n = p.field_n;
  d.x0 = (R)grid_x0(p);
  d.y0 = (R)grid_y0(p);
  d.spacing = (R)p.spacing;
  // gpu_process.py:1165-1171 (dt - 1 formed in the value type)
  const R dt = (R)p.deforming_par;
  if (dt > R(0) && dt < R(1)) {
    d.warp_a = 1;
    d.warp_b = 1;
    d.dt_a = dt - R(1);
    d.dt_b = dt;
  } else if (dt == R(1)) {
    d.warp_a = 0;
    d.warp_b = 1;
    d.dt_b = R(1);
  } else {
    d.warp_a = 1;
    d.warp_b = 0;
    d.dt_a = R(-1);
  }
  d.out_a = oa;
  d.out_b = ob;
  return d;
}

template <typename R>
GridInterpArgsT<R> interp_args(const piv_pass_desc& from, const piv_pass_desc& to, const R* ui, const R* vi, R* uo,
                               R* vo) {
  GridInterpArgsT<R> g{};
  g.u_in = ui;
  g.v_in = vi;
  g.m_in = from.field_m;
  g.n_in = from.field_n;
  g.x0_in = (R)grid_x0(from);
  g.y0_in = (R)grid_y0(from);
  g.sp_in = (R)from.spacing;
  g.m_out = to.field_m;
  g.n_out = to.field_n;
  g.x0_out = (R)grid_x0(to);
  g.y0_out = (R)grid_y0(to);
  g.sp_out = (R)to.spacing;
  g.u_out = uo;
  g.v_out = vo;
  return g;
}

}  // namespace

// Field arrays hold the plan's value type: float, or double for FP64 plans
// (the executor's templates cast them; FP32-only paths use them as float).
struct PassBuffers {
  piv_pass_desc d{};
  int N = 0;
  float* u = nullptr;
  float* v = nullptr;
  float* up = nullptr;
  float* vp = nullptr;
  uint8_t* flags = nullptr;
  float* s2n = nullptr;         // when the pass has an S2N test
  SmoothTables tab;
  double p_min = 0, p_max = 0;
  uint8_t* fmask = nullptr;     // field mask (piv_plan_set_mask) or null
  int n_masked = 0;             // masked vectors: > 0 selects weighted smoothing
  int slow_gcv = 0;             // mean weight <= 0.9
};

// Work buffers of the validation / replacement loop (field-sized).
struct LoopBufs {
  void* u0 = nullptr;         // originals for the unresolved reset (value type)
  void* v0 = nullptr;
  void* tu = nullptr;         // inputs of a simultaneous (Jacobi) replacement
  void* tv = nullptr;
  uint8_t* wf[2] = {nullptr, nullptr};  // flags after each revalidation (ping-pong)
  uint8_t* nanf = nullptr;    // unresolved vectors of the last replacement
  int32_t* it_counts = nullptr;  // PIV_MAX_REPLACING_ITERS counters per pass
  double* stats = nullptr;    // mean_u, std_u, mean_v, std_v
  ReplaceWork rw{};
};

struct piv_plan {
  int device = 0;
  piv_plan_desc desc{};
  hipStream_t stream = nullptr;
  std::vector<PassBuffers> passes;
  void* frame_a = nullptr;
  void* frame_b = nullptr;
  size_t frame_bytes = 0;
  int fp64 = 0;                 // dtype_f = float64: every field and deformed frame in double
  size_t esize = 4;             // bytes per value
  void* def_a = nullptr;        // deformed frames (value type)
  void* def_b = nullptr;
  double* twiddle[11] = {};     // FP64 FFT twiddles per log2(F) (F = 16..1024)
  void* corr_scratch = nullptr;  // plane slabs: FP64 F >= 128, FP32 F >= 256
  int32_t* counters = nullptr;  // 2 per pass
  double* p_out = nullptr;      // 1 per pass
  uint8_t* val_final = nullptr;
  void* t0 = nullptr;           // smoothing transforms (value type)
  void* t1 = nullptr;
  void* smooth_state = nullptr;
  ReplaceWork rw{};
  LoopBufs lb{};                // shares rw
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  int runs = 0;
  // learned predictor (piv_plan_set_predictor)
  piv_cnn* net = nullptr;
  int net_first = 0;            // first pass run on the learned path (its predictor = the network's field)
  float max_reject = 0.05f;
  piv_pass_desc net_grid{};     // the network's output nodes as a pass grid
  uint16_t* net_in = nullptr;   // network input of the current pair
  void* net_scratch = nullptr;
  hipGraph_t lgraph = nullptr;
  hipGraphExec_t lgraph_exec = nullptr;
  hipEvent_t inputs_ready = nullptr;  // null-stream work that produced device-resident frames
  hipEvent_t wait_next = nullptr;  // the next piv_run waits for this event (piv_plan_wait_event)
  const void* ext_a = nullptr;  // caller-owned device frames read in place (batches)
  const void* ext_b = nullptr;
  int lruns = 0;
  int32_t last_fallback = 0;
  int32_t last_rejected = 0;
  // learned runs queued by piv_run_async: their outputs and frames, and the
  // rejection counts copied to pinned host slots on the stream; the
  // fallbacks run in piv_plan_synchronize
  struct Pending {
    const void* fa;
    const void* fb;
    void* u;
    void* v;
    uint8_t* vi;
    uint8_t* vf;
  };
  std::vector<Pending> pending;
  int32_t* rej_host = nullptr;  // kMaxPendingLearned slots
  int64_t learned_runs = 0, learned_fallbacks = 0;
  // frame mask (piv_plan_set_mask)
  uint8_t* frame_mask = nullptr;
  MaskedSmoothBufs msb{};
  int32_t* smooth_iters = nullptr;  // 1 per pass
};

namespace {

const void* src_a(const piv_plan* P) { return P->ext_a ? P->ext_a : P->frame_a; }
const void* src_b(const piv_plan* P) { return P->ext_b ? P->ext_b : P->frame_b; }

template <typename R>
R* as(void* p) {
  return reinterpret_cast<R*>(p);
}

CorrArgs64 corr_args64_for(const void* fa, const void* fb, int dtype, int ht, int wd, const piv_pass_desc& p,
                           const double* twiddle, double2* scratch) {
  CorrArgs64 a{};
  a.frame_a = fa;
  a.frame_b = fb;
  a.in_dtype = dtype;
  a.ht = ht;
  a.wd = wd;
  a.ss = p.search_size;
  a.fsize = p.fft_size;
  a.spacing = p.spacing;
  a.x_off = p.x_offset;
  a.y_off = p.y_offset;
  a.m = p.field_m;
  a.n = p.field_n;
  a.normalize = p.normalize;
  a.subpixel = p.subpixel;
  a.s2n_method = p.s2n_method;
  a.s2n_size = p.s2n_size;
  a.win_pad = p.window_size > 0 ? (p.search_size - p.window_size) / 2 : 0;
  a.twiddle = twiddle;
  a.scratch = scratch;
  return a;
}

CorrArgs64 corr_args64(const piv_plan* P, const void* fa, const void* fb, int dtype, const piv_pass_desc& p) {
  int lg = 0;
  while ((1 << lg) < p.fft_size) ++lg;
  return corr_args64_for(fa, fb, dtype, P->desc.height, P->desc.width, p, P->twiddle[lg],
                         (double2*)P->corr_scratch);
}

// Pass k > 0 (or the learned path's single pass): deform the original
// frames by the total predictor B.up/B.vp (gpu_process.py:1113-1114, Q13)
// and correlate the residual, adding the predictor back (:628-632).
template <typename R>
int deform_and_correlate(piv_plan* P, PassBuffers& B, hipStream_t s) {
  const piv_plan_desc& D = P->desc;
  DeformArgsT<R> da = deform_args<R>(src_a(P), src_b(P), D.in_dtype, D.height, D.width, B.d, as<R>(B.up),
                                     as<R>(B.vp), as<R>(P->def_a), as<R>(P->def_b));
  da.mask = P->frame_mask;
  HIP_TRY(launch_deform(da, s));
  if constexpr (sizeof(R) == 8) {
    CorrArgs64 a = corr_args64(P, P->def_a, P->def_b, kF64, B.d);
    a.up = as<double>(B.up);
    a.vp = as<double>(B.vp);
    a.u = as<double>(B.u);
    a.v = as<double>(B.v);
    a.s2n = as<double>(B.s2n);
    HIP_TRY(launch_correlate64(a, s));
  } else {
    CorrArgs a = corr_args(P->def_a, P->def_b, kF32, D.height, D.width, B.d);
    a.scratch = P->corr_scratch;
    a.up = B.up;
    a.vp = B.vp;
    a.u = B.u;
    a.v = B.v;
    a.s2n = B.s2n;
    HIP_TRY(launch_correlate(a, s));
  }
  if (B.fmask) HIP_TRY(launch_apply_field_mask(as<R>(B.u), as<R>(B.v), as<R>(B.s2n), B.fmask, B.N, s));
  return PIV_OK;
}

// Pass 0: correlate the raw (masked) frames.
template <typename R>
int correlate_first(piv_plan* P, PassBuffers& B, hipStream_t s) {
  const piv_plan_desc& D = P->desc;
  if constexpr (sizeof(R) == 8) {
    CorrArgs64 a = corr_args64(P, src_a(P), src_b(P), D.in_dtype, B.d);
    a.u = as<double>(B.u);
    a.v = as<double>(B.v);
    a.s2n = as<double>(B.s2n);
    HIP_TRY(launch_correlate64(a, s));
  } else {
    CorrArgs a = corr_args(src_a(P), src_b(P), D.in_dtype, D.height, D.width, B.d);
    a.scratch = P->corr_scratch;
    a.u = B.u;
    a.v = B.v;
    a.s2n = B.s2n;
    HIP_TRY(launch_correlate(a, s));
  }
  if (B.fmask) HIP_TRY(launch_apply_field_mask(as<R>(B.u), as<R>(B.v), as<R>(B.s2n), B.fmask, B.N, s));
  return PIV_OK;
}

template <typename R>
ValidateArgsT<R> validate_args(const piv_pass_desc& d, const R* u, const R* v, int m, int n, const R* s2n,
                               const double* stats, const uint8_t* only, const uint8_t* extra, uint8_t* flags,
                               int32_t* count, const uint8_t* fmask = nullptr) {
  ValidateArgsT<R> a{};
  a.u = u;
  a.v = v;
  a.m = m;
  a.n = n;
  a.size = validation_size(d);
  a.s2n = s2n;
  a.s2n_tol = s2n ? (R)d.s2n_tol : (R)NAN;
  a.median_tol = (R)d.median_tol;
  a.mad_tol = (R)d.mad_tol;
  a.mean_tol = (R)d.mean_tol;
  a.rms_tol = (R)d.rms_tol;
  a.rms = stats;
  a.only = only;
  a.extra = extra;
  a.flags = flags;
  a.count = count;
  a.fmask = fmask;
  return a;
}

// ValidationGPU.__call__ (gpu_validation.py:84-118): every enabled test.
template <typename R>
int run_validate(const piv_pass_desc& d, const R* u, const R* v, int m, int n, const R* s2n, double* stats,
                 const uint8_t* only, const uint8_t* extra, uint8_t* flags, int32_t* count, hipStream_t s,
                 const uint8_t* fmask = nullptr) {
  if (!std::isnan(d.rms_tol)) HIP_TRY(launch_field_stats(u, v, m * n, stats, s));
  HIP_TRY(launch_validate(validate_args(d, u, v, m, n, s2n, stats, only, extra, flags, count, fmask), s));
  return PIV_OK;
}

// PIVGPU.replace_outliers (gpu_process.py:788-834) on the vectors flagged in
// flags0 (count0 on the device): num_replacing_iters iterations of the
// pass's method; iteration 1 leaves the flagged neighbours out, later ones
// use them unless revalidating (:824-826); revalidation resets unresolved
// vectors and re-tests only the replaced ones (:809-822); without it the
// unresolved ones are reset at the end (:829-832).  it_counts: this pass's
// PIV_MAX_REPLACING_ITERS revalidation counters, zero on entry.
// No host synchronisation: an empty flag set turns every step into a no-op.
template <typename R>
int replace_loop(R* u, R* v, const uint8_t* flags0, int32_t* count0, int m, int n, const piv_pass_desc& d,
                 const R* s2n, LoopBufs& L, int32_t* it_counts, hipStream_t s, const uint8_t* fmask = nullptr) {
  const int N = m * n;
  const int iters = d.num_replacing_iters;
  if (iters <= 0) return PIV_OK;
  const bool spring = d.replacing_method == PIV_SPRING;
  const bool reval = d.revalidate != 0;
  R* u0 = as<R>(L.u0);
  R* v0 = as<R>(L.v0);
  R* tu = as<R>(L.tu);
  R* tv = as<R>(L.tv);
  if (!spring || reval) {
    HIP_TRY(launch_copy(u0, (const R*)u, N, s));
    HIP_TRY(launch_copy(v0, (const R*)v, N, s));
  }
  const uint8_t* cur = flags0;
  int32_t* cur_count = count0;
  for (int it = 0; it < iters; ++it) {
    if (spring) {
      HIP_TRY(launch_spring_replace(u, v, cur, cur_count, m, n, replacing_size(d), L.rw, s));
    } else {
      HIP_TRY(launch_copy(tu, (const R*)u, N, s));
      HIP_TRY(launch_copy(tv, (const R*)v, N, s));
      HIP_TRY(launch_neighbour_replace((const R*)tu, (const R*)tv, u, v, cur, m, n, replacing_size(d), d.replacing_method,
                                       (it == 0 || reval) ? 1 : 0, s));
    }
    if (reval) {
      HIP_TRY(launch_reset_unresolved(u, v, (const R*)u0, (const R*)v0, cur, L.nanf, N, s));
      uint8_t* nxt = L.wf[it & 1];
      int32_t* nxt_count = it_counts + it;
      int r = run_validate<R>(d, u, v, m, n, s2n, L.stats, cur, L.nanf, nxt, nxt_count, s, fmask);
      if (r) return r;
      cur = nxt;
      cur_count = nxt_count;
    }
  }
  if (!reval && !spring) HIP_TRY(launch_reset_unresolved(u, v, (const R*)u0, (const R*)v0, cur, nullptr, N, s));
  return PIV_OK;
}

// Validation, replacement, last-pass revalidation and smoothing of pass k
// (gpu_process.py:635-639, 744-845).
template <typename R>
int validate_and_smooth(piv_plan* P, int k, bool last, hipStream_t s) {
  PassBuffers& B = P->passes[k];
  const int m = B.d.field_m, n = B.d.field_n;
  R* u = as<R>(B.u);
  R* v = as<R>(B.v);
  const R* s2n = as<R>(B.s2n);
  if (B.d.validate) {
    int r = run_validate<R>(B.d, u, v, m, n, s2n, P->lb.stats, nullptr, nullptr, B.flags, P->counters + 2 * k, s,
                            B.fmask);
    if (!r)
      r = replace_loop<R>(u, v, B.flags, P->counters + 2 * k, m, n, B.d, s2n, P->lb,
                          P->lb.it_counts + (size_t)k * PIV_MAX_REPLACING_ITERS, s, B.fmask);
    if (!r && last)
      r = run_validate<R>(B.d, u, v, m, n, s2n, P->lb.stats, nullptr, nullptr, P->val_final, P->counters + 2 * k + 1,
                          s, B.fmask);
    if (r) return r;
  }
  if (!last && B.d.smooth && B.n_masked > 0) {
    const int autos = !(B.d.smoothing_par > 0);
    SmoothnMaskedWorkT<R> w = masked_work<R>(B.tab, B.fmask, P->msb, P->t0, P->t1, P->p_out + k,
                                             P->smooth_iters + k, B.slow_gcv);
    HIP_TRY(launch_smoothn_masked(u, v, m, n, B.p_min, B.p_max, autos ? 0.0 : log10(B.d.smoothing_par), autos, w,
                                  s));
  } else if (!last && B.d.smooth) {
    SmoothnWorkT<R> w{};
    w.cos_m = (const R*)B.tab.cos_m;
    w.cos_n = (const R*)B.tab.cos_n;
    w.lambda = B.tab.lambda;
    w.t0 = as<R>(P->t0);
    w.t1 = as<R>(P->t1);
    w.p_out = P->p_out + k;
    w.state = P->smooth_state;
    const int autos = !(B.d.smoothing_par > 0);
    HIP_TRY(launch_smoothn(u, v, B.d.field_m, B.d.field_n, B.p_min, B.p_max,
                           autos ? 0.0 : log10(B.d.smoothing_par), autos, w, s));
  }
  return PIV_OK;
}

// Pass k with the predictor from pass k-1 (gpu_process.py:620-625).
template <typename R>
int predictor_from_previous(piv_plan* P, int k, hipStream_t s) {
  PassBuffers& A = P->passes[k - 1];
  PassBuffers& B = P->passes[k];
  if (A.d.search_size != B.d.search_size) {
    HIP_TRY(launch_grid_interp(
        interp_args<R>(A.d, B.d, as<R>(A.u), as<R>(A.v), as<R>(B.up), as<R>(B.vp)), s));
  } else {
    HIP_TRY(launch_copy(as<R>(B.up), (const R*)as<R>(A.u), B.N, s));
    HIP_TRY(launch_copy(as<R>(B.vp), (const R*)as<R>(A.v), B.N, s));
  }
  return PIV_OK;
}

// Passes first..np-1; pass `first` takes its predictor from B.up/B.vp as
// already written (first > 0) or correlates the raw frames (first == 0).
template <typename R>
int enqueue_passes_t(piv_plan* P, int first) {
  hipStream_t s = P->stream;
  const piv_plan_desc& D = P->desc;
  const int np = (int)P->passes.size();
  for (int k = first; k < np; ++k) {
    PassBuffers& B = P->passes[k];
    const bool last = k == np - 1;
    int r = PIV_OK;
    if (k == 0) {
      r = correlate_first<R>(P, B, s);
    } else {
      if (k > first) r = predictor_from_previous<R>(P, k, s);
      if (!r) r = deform_and_correlate<R>(P, B, s);
    }
    if (!r) r = validate_and_smooth<R>(P, k, last, s);
    if (r) return r;
  }
  PassBuffers& L = P->passes.back();
  if constexpr (sizeof(R) == 8) {
    // (u * scaling_par) / dt in double (gpu_process.py:642-643, 847-849)
    if (D.scaling_par != 1.0 || D.dt != 1.0)
      HIP_TRY(launch_scale_div(as<double>(L.u), as<double>(L.v), L.N, D.scaling_par, D.dt, s));
  } else if (D.out_scale != 1.0f) {
    HIP_TRY(launch_scale(as<R>(L.u), as<R>(L.v), L.N, (R)D.out_scale, s));
  }
  return PIV_OK;
}

int enqueue_passes(piv_plan* P, int first) {
  return P->fp64 ? enqueue_passes_t<double>(P, first) : enqueue_passes_t<float>(P, first);
}

// The run's counters (outlier counts, replacement iteration counts) are
// cleared by piv_run's input launch before the graph starts.
int enqueue(piv_plan* P) {
  if (P->frame_mask)
    HIP_TRY(launch_mask_frames(P->frame_a, P->frame_b, P->desc.in_dtype, P->frame_mask,
                               (size_t)P->desc.height * P->desc.width, P->stream));
  return enqueue_passes(P, 0);
}

// Learned path: the network's field, interpolated to the grid of pass
// net_first, is that pass's deformation predictor; earlier passes do not run.
int enqueue_learned(piv_plan* P) {
  hipStream_t s = P->stream;
  const piv_plan_desc& D = P->desc;
  PassBuffers& B = P->passes[P->net_first];
  HIP_TRY(launch_prepare_input(src_a(P), src_b(P), D.in_dtype, D.height, D.width, kCnnFactor, P->net_scratch,
                               P->net_in, s));
  int r = cnn_enqueue_forward(P->net, P->net_in, 1, s);
  if (r) return r;
  const float* o = cnn_output(P->net);
  const CnnInfo ni = cnn_info(P->net);
  HIP_TRY(launch_grid_interp(interp_args(P->net_grid, B.d, o, o + (size_t)ni.out_h * ni.out_w, B.up, B.vp), s));
  return enqueue_passes(P, P->net_first);
}

// Runs an enqueue function eagerly on its first use and as a captured graph
// afterwards (when the plan asks for graphs).
int run_captured(piv_plan* P, int (*fn)(piv_plan*), hipGraph_t* graph, hipGraphExec_t* exec, int* runs) {
  hipStream_t s = P->stream;
  if (P->desc.use_graph && *runs > 0) {
    if (!*exec) {
      HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      int r = fn(P);
      hipGraph_t g = nullptr;
      hipError_t e = hipStreamEndCapture(s, &g);
      if (r) {
        if (g) (void)hipGraphDestroy(g);
        return r;
      }
      if (e != hipSuccess) return fail(PIV_E_HIP, "graph capture failed: %s", hipGetErrorString(e));
      *graph = g;
      HIP_TRY(hipGraphInstantiate(exec, *graph, nullptr, nullptr, 0));
    }
    HIP_TRY(hipGraphLaunch(*exec, s));
  } else {
    int r = fn(P);
    if (r) return r;
  }
  (*runs)++;
  return PIV_OK;
}

void drop_learned_graph(piv_plan* P) {
  if (P->lgraph_exec) (void)hipGraphExecDestroy(P->lgraph_exec);
  if (P->lgraph) (void)hipGraphDestroy(P->lgraph);
  P->lgraph_exec = nullptr;
  P->lgraph = nullptr;
  P->lruns = 0;
}

void release(piv_plan* P) {
  if (!P) return;
  (void)hipSetDevice(P->device);
  if (P->graph_exec) (void)hipGraphExecDestroy(P->graph_exec);
  if (P->graph) (void)hipGraphDestroy(P->graph);
  drop_learned_graph(P);
  if (P->rej_host) (void)hipHostFree(P->rej_host);
  if (P->net_in) (void)hipFree(P->net_in);
  if (P->net_scratch) (void)hipFree(P->net_scratch);
  for (auto& B : P->passes) {
    for (void* p : {(void*)B.u, (void*)B.v, (void*)B.up, (void*)B.vp, (void*)B.flags, (void*)B.s2n, (void*)B.fmask})
      if (p) (void)hipFree(p);
    free_tables(&B.tab);
  }
  free_masked_smooth(&P->msb);
  if (P->frame_mask) (void)hipFree(P->frame_mask);
  if (P->smooth_iters) (void)hipFree(P->smooth_iters);
  for (void* p : {(void*)P->lb.u0, (void*)P->lb.v0, (void*)P->lb.tu, (void*)P->lb.tv, (void*)P->lb.wf[0],
                  (void*)P->lb.wf[1], (void*)P->lb.nanf, (void*)P->lb.it_counts, (void*)P->lb.stats})
    if (p) (void)hipFree(p);
  for (void* p : {P->frame_a, P->frame_b, (void*)P->def_a, (void*)P->def_b, (void*)P->counters, (void*)P->p_out,
                  (void*)P->val_final, (void*)P->t0, (void*)P->t1, (void*)P->rw.slot, (void*)P->rw.list,
                  (void*)P->rw.work, P->smooth_state, (void*)P->corr_scratch})
    if (p) (void)hipFree(p);
  for (double* t : P->twiddle)
    if (t) (void)hipFree(t);
  if (P->inputs_ready) (void)hipEventDestroy(P->inputs_ready);
  if (P->stream) (void)hipStreamDestroy(P->stream);
  delete P;
}

// The replacement loop of piv_stage_replace on caller-provided work
// buffers: the incoming flag count, replace_loop, and the final flags of a
// revalidating loop.  Counters are cleared by fill kernels (no memset), so
// the sequence can also be captured into a graph.
int enqueue_stage_replace(float* u, float* v, uint8_t* flags, int m, int n, const piv_pass_desc& pass,
                          const float* s2n, LoopBufs& L, int32_t* count0, hipStream_t s) {
  const size_t N = (size_t)m * n;
  hipError_t e = launch_fill_i32(L.it_counts, 0, PIV_MAX_REPLACING_ITERS, s);
  if (e == hipSuccess) e = launch_fill_i32(count0, 0, 1, s);
  // the count of the incoming flags (the validation kernel's counter)
  if (e == hipSuccess) {
    piv_pass_desc none = pass;
    none.median_tol = none.mad_tol = none.mean_tol = none.rms_tol = none.s2n_tol = NAN;
    e = launch_validate(validate_args<float>(none, u, v, m, n, nullptr, nullptr, nullptr, flags, L.wf[1], count0), s);
  }
  if (e != hipSuccess) return fail(PIV_E_HIP, "replace: %s", hipGetErrorString(e));
  int r = replace_loop(u, v, L.wf[1], count0, m, n, pass, s2n, L, L.it_counts, s);
  if (!r && pass.revalidate && pass.num_replacing_iters > 0) {
    // final flags of the loop: the last revalidation's output
    const uint8_t* last = L.wf[(pass.num_replacing_iters - 1) & 1];
    e = hipMemcpyAsync(flags, last, N, hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) r = fail(PIV_E_HIP, "replace: %s", hipGetErrorString(e));
  }
  return r;
}

}  // namespace

// Fixed workspace of piv_stage_replace_noalloc, one copy per GPU for the life
// of the process: u0, v0, tu, tv (4 B per vector each), two revalidation flag
// buffers and the unresolved flags (1 B each), the spring slot and list
// (4 B each) and its 8 doubles per vector, then the small counters -- each
// part on a 256-byte boundary.
namespace {
constexpr size_t kNoAllocVec = PIV_NOALLOC_MAX_VECTORS;
constexpr size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }
constexpr size_t kNoAllocBytes =
    4 * align256(4 * kNoAllocVec) + 3 * align256(kNoAllocVec) + 2 * align256(4 * kNoAllocVec) +
    align256(64 * kNoAllocVec) + align256(4 * PIV_MAX_REPLACING_ITERS) + align256(32) + align256(4);
}  // namespace

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxDevices] = {};

// The device address of this GPU's copy of `scratch`, looked up once per
// device with that device current (the array's name is only a host-side
// handle: passing it to a kernel would fault).
int noalloc_scratch(int dev, unsigned char** out) {
  if (dev < 0 || dev >= kMaxDevices) return fail(PIV_E_INVALID, "device %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(PIV_E_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_scratch_ptr[dev] = (unsigned char*)p;
  }
  *out = g_scratch_ptr[dev];
  return PIV_OK;
}

// Work buffers for n vectors (n <= PIV_NOALLOC_MAX_VECTORS) carved out of the
// fixed workspace in the order of the size formula above.
void noalloc_layout(unsigned char* base, size_t n, LoopBufs* L, int32_t** count0) {
  (void)n;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    unsigned char* p = base + off;
    off += align256(bytes);
    return (void*)p;
  };
  L->u0 = take(4 * kNoAllocVec);
  L->v0 = take(4 * kNoAllocVec);
  L->tu = take(4 * kNoAllocVec);
  L->tv = take(4 * kNoAllocVec);
  L->wf[0] = (uint8_t*)take(kNoAllocVec);
  L->wf[1] = (uint8_t*)take(kNoAllocVec);
  L->nanf = (uint8_t*)take(kNoAllocVec);
  L->rw.slot = (int32_t*)take(4 * kNoAllocVec);
  L->rw.list = (int32_t*)take(4 * kNoAllocVec);
  L->rw.work = (double*)take(64 * kNoAllocVec);
  L->it_counts = (int32_t*)take(4 * PIV_MAX_REPLACING_ITERS);
  L->stats = (double*)take(32);
  *count0 = (int32_t*)take(4);
}

}  // namespace

extern "C" {

const char* piv_last_error(void) { return g_err.c_str(); }

int piv_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int piv_plan_create(int device, const piv_plan_desc* desc, piv_plan** out) {
  if (!desc || !out) return fail(PIV_E_INVALID, "null argument");
  *out = nullptr;
  if (piv_device_count() <= 0) return fail(PIV_E_NO_DEVICE, "no HIP device visible");
  const piv_plan_desc& D = *desc;
  if (D.n_passes < 1 || D.n_passes > PIV_MAX_PASSES) return fail(PIV_E_INVALID, "n_passes out of range");
  if (D.height < 2 || D.width < 2) return fail(PIV_E_INVALID, "frame too small");
  if (D.in_dtype < PIV_U8 || D.in_dtype > PIV_F64) return fail(PIV_E_INVALID, "in_dtype");
  if (D.precision != PIV_FP32 && D.precision != PIV_FP64) return fail(PIV_E_INVALID, "precision");
  if (D.in_dtype == PIV_F64 && D.precision != PIV_FP64)
    return fail(PIV_E_INVALID, "float64 frames need an FP64 plan (convert to float32 for FP32 processing)");
  for (int k = 0; k < D.n_passes; ++k) {
    int r = check_pass(D.passes[k], D.height, D.width);
    if (r) return r;
    if (k > 0 && (D.passes[k].field_m < 2 || D.passes[k].field_n < 2 || D.passes[k - 1].field_m < 2 ||
                  D.passes[k - 1].field_n < 2))
      return fail(PIV_E_INVALID, "multi-pass schedules need fields of at least 2 x 2");
  }
  HIP_TRY(hipSetDevice(device));
  piv_plan* P = new piv_plan();
  P->device = device;
  P->desc = D;
  P->fp64 = D.precision == PIV_FP64;
  P->esize = P->fp64 ? 8 : 4;
  const size_t es = P->esize;
  auto bail = [&](int code) {
    release(P);
    return code;
  };
  if (hipStreamCreateWithFlags(&P->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&P->inputs_ready, hipEventDisableTiming) != hipSuccess)
    return bail(fail(PIV_E_HIP, "hipStreamCreate/hipEventCreate failed"));
  const size_t npx = (size_t)D.height * D.width;
  P->frame_bytes = npx * dtype_size(D.in_dtype);
  int maxN = 0;
  size_t scratch_bytes = 0;
  auto dmalloc = [&](void** p, size_t bytes) -> bool { return hipMalloc(p, bytes ? bytes : 16) == hipSuccess; };
  bool ok = dmalloc(&P->frame_a, P->frame_bytes) && dmalloc(&P->frame_b, P->frame_bytes);
  if (D.n_passes > 1) ok = ok && dmalloc((void**)&P->def_a, npx * es) && dmalloc((void**)&P->def_b, npx * es);
  ok = ok && dmalloc((void**)&P->counters, 2 * D.n_passes * sizeof(int32_t)) &&
       dmalloc((void**)&P->p_out, D.n_passes * sizeof(double));
  if (!ok) return bail(fail(PIV_E_NOMEM, "device allocation failed"));
  {
    std::vector<double> nanv(D.n_passes, NAN);
    if (hipMemcpy(P->p_out, nanv.data(), nanv.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
      return bail(fail(PIV_E_HIP, "hipMemcpy failed"));
  }
  P->passes.resize(D.n_passes);
  for (int k = 0; k < D.n_passes; ++k) {
    PassBuffers& B = P->passes[k];
    B.d = D.passes[k];
    B.N = B.d.field_m * B.d.field_n;
    if (B.N > maxN) maxN = B.N;
    ok = dmalloc((void**)&B.u, B.N * es) && dmalloc((void**)&B.v, B.N * es) && dmalloc((void**)&B.up, B.N * es) &&
         dmalloc((void**)&B.vp, B.N * es) && dmalloc((void**)&B.flags, B.N);
    if (ok && needs_s2n(B.d)) ok = dmalloc((void**)&B.s2n, B.N * es);
    if (!ok) return bail(fail(PIV_E_NOMEM, "device allocation failed"));
    if (hipMemset(B.flags, 0, B.N) != hipSuccess) return bail(fail(PIV_E_HIP, "hipMemset failed"));
    if (P->fp64) {
      int lg = 0;
      while ((1 << lg) < B.d.fft_size) ++lg;
      if (!P->twiddle[lg]) {
        int r = upload((void**)&P->twiddle[lg], fft_twiddles(B.d.fft_size));
        if (r) return bail(r);
      }
    }
    // one scratch for every pass: the largest any of them needs
    scratch_bytes = std::max(scratch_bytes, P->fp64 ? correlate64_scratch_bytes(B.d.fft_size)
                                                    : correlate_big_scratch_bytes(B.d.fft_size, 4, B.N));
    if (k < D.n_passes - 1 && B.d.smooth) {
      int r = make_tables(B.d.field_m, B.d.field_n, &B.tab, P->fp64 != 0);
      if (r) return bail(r);
      p_bounds(B.d.field_m, B.d.field_n, &B.p_min, &B.p_max);
    }
  }
  if (scratch_bytes && !dmalloc(&P->corr_scratch, scratch_bytes))
    return bail(fail(PIV_E_NOMEM, "device allocation failed"));
  const PassBuffers& L = P->passes.back();
  ok = dmalloc((void**)&P->val_final, L.N) && dmalloc((void**)&P->t0, 2ull * maxN * es) &&
       dmalloc((void**)&P->t1, 2ull * maxN * es) && dmalloc((void**)&P->rw.slot, maxN * 4) &&
       dmalloc((void**)&P->rw.list, maxN * 4) && dmalloc((void**)&P->rw.work, 8ull * maxN * 8);
  ok = ok && dmalloc(&P->smooth_state, smoothn_state_bytes());
  {
    LoopBufs& lb = P->lb;
    ok = ok && dmalloc((void**)&lb.u0, maxN * es) && dmalloc((void**)&lb.v0, maxN * es) &&
         dmalloc((void**)&lb.tu, maxN * es) && dmalloc((void**)&lb.tv, maxN * es) && dmalloc((void**)&lb.wf[0], maxN) &&
         dmalloc((void**)&lb.wf[1], maxN) && dmalloc((void**)&lb.nanf, maxN) &&
         dmalloc((void**)&lb.it_counts, (size_t)D.n_passes * PIV_MAX_REPLACING_ITERS * 4) && dmalloc((void**)&lb.stats, 4 * 8);
    lb.rw = P->rw;
  }
  if (!ok) return bail(fail(PIV_E_NOMEM, "device allocation failed"));
  if (hipMemset(P->val_final, 0, L.N) != hipSuccess || hipMemset(P->smooth_state, 0, smoothn_state_bytes()) != hipSuccess)
    return bail(fail(PIV_E_HIP, "hipMemset failed"));
  *out = P;
  return PIV_OK;
}

void piv_plan_destroy(piv_plan* plan) { release(plan); }

// Learned runs queued per plan between synchronisations (the pinned slots of
// their rejection counts).
constexpr int kMaxPendingLearned = 1024;

static int copy_outputs(piv_plan* P, void* u, void* v, uint8_t* val_init, uint8_t* val_final, int outputs_on_device);
static int resolve_pending(piv_plan* P);

static int run_impl(piv_plan* P, const void* frame_a, const void* frame_b, int frames_on_device, void* u, void* v,
                    uint8_t* val_init, uint8_t* val_final, int outputs_on_device, bool sync, bool defer = false) {
  if (!P || !frame_a || !frame_b) return fail(PIV_E_INVALID, "null argument");
  HIP_TRY(hipSetDevice(P->device));
  hipStream_t s = P->stream;
  const hipMemcpyKind in_kind = frames_on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
  if (frames_on_device) {
    // device frames may come from work queued on the null stream, which a
    // non-blocking plan stream does not wait for by itself
    HIP_TRY(hipEventRecord(P->inputs_ready, nullptr));
    HIP_TRY(hipStreamWaitEvent(s, P->inputs_ready, 0));
  }
  if (P->wait_next) {
    HIP_TRY(hipStreamWaitEvent(s, P->wait_next, 0));
    P->wait_next = nullptr;
  }
  {
    // frames in, and the run's counters cleared, in one launch
    CopyList cl{};
    int nc = 0;
    cl.c[nc++] = CopySpec{P->counters, nullptr, 2 * P->passes.size() * sizeof(int32_t)};
    cl.c[nc++] = CopySpec{P->lb.it_counts, nullptr, P->passes.size() * PIV_MAX_REPLACING_ITERS * sizeof(int32_t)};
    if (frames_on_device) {
      cl.c[nc++] = CopySpec{P->frame_a, frame_a, P->frame_bytes};
      cl.c[nc++] = CopySpec{P->frame_b, frame_b, P->frame_bytes};
    } else {
      HIP_TRY(hipMemcpyAsync(P->frame_a, frame_a, P->frame_bytes, in_kind, s));
      HIP_TRY(hipMemcpyAsync(P->frame_b, frame_b, P->frame_bytes, in_kind, s));
    }
    HIP_TRY(launch_copy_multi(cl, nc, s));
  }
  // The first run of each path is eager (it also sets kernel attributes);
  // later runs replay a captured graph when requested.
  P->last_fallback = 0;
  P->last_rejected = 0;
  if (P->net && defer) {
    // queued learned run: the rejection count goes to a pinned slot on the
    // stream; piv_plan_synchronize runs the fallbacks
    int r = run_captured(P, enqueue_learned, &P->lgraph, &P->lgraph_exec, &P->lruns);
    if (r) return r;
    const size_t slot = P->pending.size();
    HIP_TRY(hipMemcpyAsync(P->rej_host + slot, P->counters + 2 * (P->passes.size() - 1), sizeof(int32_t),
                           hipMemcpyDeviceToHost, s));
    P->pending.push_back(piv_plan::Pending{frame_a, frame_b, u, v, val_init, val_final});
    return copy_outputs(P, u, v, val_init, val_final, outputs_on_device);
  }
  if (P->net) {
    int r = run_captured(P, enqueue_learned, &P->lgraph, &P->lgraph_exec, &P->lruns);
    if (r) return r;
    const PassBuffers& L = P->passes.back();
    int32_t rejected = 0;
    HIP_TRY(hipMemcpyAsync(&rejected, P->counters + 2 * (P->passes.size() - 1), sizeof(int32_t),
                           hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    P->last_rejected = rejected;
    P->learned_runs++;
    if ((double)rejected > (double)P->max_reject * L.N) {
      P->last_fallback = 1;
      P->learned_fallbacks++;
      CopyList cl{};  // counters of the full-schedule run start at zero
      cl.c[0] = CopySpec{P->counters, nullptr, 2 * P->passes.size() * sizeof(int32_t)};
      cl.c[1] = CopySpec{P->lb.it_counts, nullptr, P->passes.size() * PIV_MAX_REPLACING_ITERS * sizeof(int32_t)};
      HIP_TRY(launch_copy_multi(cl, 2, s));
      r = run_captured(P, enqueue, &P->graph, &P->graph_exec, &P->runs);
      if (r) return r;
    }
  } else {
    int r = run_captured(P, enqueue, &P->graph, &P->graph_exec, &P->runs);
    if (r) return r;
  }
  int r = copy_outputs(P, u, v, val_init, val_final, outputs_on_device);
  if (r) return r;
  if (sync) HIP_TRY(hipStreamSynchronize(s));
  return PIV_OK;
}

static int copy_outputs(piv_plan* P, void* u, void* v, uint8_t* val_init, uint8_t* val_final, int outputs_on_device) {
  hipStream_t s = P->stream;
  const PassBuffers& L = P->passes.back();
  if (outputs_on_device) {  // one launch for all requested outputs
    CopyList cl{};
    int nc = 0;
    if (u) cl.c[nc++] = CopySpec{u, L.u, L.N * P->esize};
    if (v) cl.c[nc++] = CopySpec{v, L.v, L.N * P->esize};
    if (val_init) cl.c[nc++] = CopySpec{val_init, L.flags, (size_t)L.N};
    if (val_final) cl.c[nc++] = CopySpec{val_final, P->val_final, (size_t)L.N};
    HIP_TRY(launch_copy_multi(cl, nc, s));
  } else {
    if (u) HIP_TRY(hipMemcpyAsync(u, L.u, L.N * P->esize, hipMemcpyDeviceToHost, s));
    if (v) HIP_TRY(hipMemcpyAsync(v, L.v, L.N * P->esize, hipMemcpyDeviceToHost, s));
    if (val_init) HIP_TRY(hipMemcpyAsync(val_init, L.flags, L.N, hipMemcpyDeviceToHost, s));
    if (val_final) HIP_TRY(hipMemcpyAsync(val_final, P->val_final, L.N, hipMemcpyDeviceToHost, s));
  }
  return PIV_OK;
}

// After the stream has drained: every queued learned run whose last-pass
// test rejected more than max_reject of the vectors reruns the full schedule
// on its frames (still valid by the async contract) into its outputs.
static int resolve_pending(piv_plan* P) {
  if (P->pending.empty()) return PIV_OK;
  std::vector<piv_plan::Pending> q;
  q.swap(P->pending);
  const PassBuffers& L = P->passes.back();
  hipStream_t s = P->stream;
  for (size_t i = 0; i < q.size(); ++i) {
    const int32_t rejected = P->rej_host[i];
    P->learned_runs++;
    P->last_rejected = rejected;
    P->last_fallback = 0;
    if ((double)rejected > (double)P->max_reject * L.N) {
      P->last_fallback = 1;
      P->learned_fallbacks++;
      const piv_plan::Pending& e = q[i];
      CopyList cl{};
      int nc = 0;
      cl.c[nc++] = CopySpec{P->counters, nullptr, 2 * P->passes.size() * sizeof(int32_t)};
      cl.c[nc++] = CopySpec{P->lb.it_counts, nullptr, P->passes.size() * PIV_MAX_REPLACING_ITERS * sizeof(int32_t)};
      cl.c[nc++] = CopySpec{P->frame_a, e.fa, P->frame_bytes};
      cl.c[nc++] = CopySpec{P->frame_b, e.fb, P->frame_bytes};
      HIP_TRY(launch_copy_multi(cl, nc, s));
      int r = run_captured(P, enqueue, &P->graph, &P->graph_exec, &P->runs);
      if (!r) r = copy_outputs(P, e.u, e.v, e.vi, e.vf, 1);
      if (r) return r;
    }
  }
  HIP_TRY(hipStreamSynchronize(s));
  return PIV_OK;
}

int piv_run(piv_plan* P, const void* frame_a, const void* frame_b, int frames_on_device, void* u, void* v,
            uint8_t* val_init, uint8_t* val_final, int outputs_on_device) {
  if (P && !P->pending.empty()) {  // queued learned runs first, fallbacks included
    int r = piv_plan_synchronize(P);
    if (r) return r;
  }
  return run_impl(P, frame_a, frame_b, frames_on_device, u, v, val_init, val_final, outputs_on_device, true);
}

int piv_run_async(piv_plan* P, const void* frame_a, const void* frame_b, void* u, void* v, uint8_t* val_init,
                  uint8_t* val_final) {
  if (!P) return fail(PIV_E_INVALID, "null plan");
  if (P->net) {
    HIP_TRY(hipSetDevice(P->device));
    if (!P->rej_host && hipHostMalloc((void**)&P->rej_host, kMaxPendingLearned * sizeof(int32_t)) != hipSuccess) {
      P->rej_host = nullptr;
      return fail(PIV_E_NOMEM, "pinned allocation failed");
    }
    if ((int)P->pending.size() >= kMaxPendingLearned) {  // make room: resolve what is queued
      HIP_TRY(hipStreamSynchronize(P->stream));
      int r = resolve_pending(P);
      if (r) return r;
    }
    return run_impl(P, frame_a, frame_b, 1, u, v, val_init, val_final, 1, false, true);
  }
  return run_impl(P, frame_a, frame_b, 1, u, v, val_init, val_final, 1, false);
}

int piv_plan_synchronize(piv_plan* P) {
  if (!P) return fail(PIV_E_INVALID, "null plan");
  HIP_TRY(hipSetDevice(P->device));
  HIP_TRY(hipStreamSynchronize(P->stream));
  return resolve_pending(P);
}

int piv_plan_learned_counts(const piv_plan* P, int64_t* runs, int64_t* fallbacks) {
  if (!P) return fail(PIV_E_INVALID, "null plan");
  if (runs) *runs = P->learned_runs;
  if (fallbacks) *fallbacks = P->learned_fallbacks;
  return PIV_OK;
}

int piv_plan_set_predictor(piv_plan* P, piv_cnn* net, int first_pass, float max_reject_fraction) {
  if (!P) return fail(PIV_E_INVALID, "null plan");
  HIP_TRY(hipSetDevice(P->device));
  HIP_TRY(hipStreamSynchronize(P->stream));
  {
    const int r = resolve_pending(P);  // queued runs finish under the predictor they were queued with
    if (r) return r;
  }
  drop_learned_graph(P);
  if (!net) {
    P->net = nullptr;
    return PIV_OK;
  }
  if (!(max_reject_fraction >= 0.f && max_reject_fraction <= 1.f))
    return fail(PIV_E_INVALID, "max_reject_fraction must lie in [0, 1]");
  if (P->frame_mask) return fail(PIV_E_INVALID, "a learned predictor cannot be combined with a frame mask");
  if (P->fp64) return fail(PIV_E_INVALID, "the learned predictor runs FP32 plans only");
  const piv_plan_desc& D = P->desc;
  const CnnInfo ni = cnn_info(net);
  if (D.height != ni.in_h * kCnnFactor || D.width != ni.in_w * kCnnFactor)
    return fail(PIV_E_INVALID, "predictor input %d x %d does not match the frame %d x %d / %d", ni.in_h, ni.in_w,
                D.height, D.width, kCnnFactor);
  if (P->passes.size() < 2) return fail(PIV_E_INVALID, "a learned predictor needs a multi-pass schedule");
  if (first_pass < 0) first_pass = (int)P->passes.size() - 1;
  if (first_pass < 1 || first_pass >= (int)P->passes.size())
    return fail(PIV_E_INVALID, "first_pass %d outside 1..%d", first_pass, (int)P->passes.size() - 1);
  if (!P->net_in) {
    if (hipMalloc((void**)&P->net_in, (size_t)ni.in_h * ni.in_w * 2 * sizeof(uint16_t)) != hipSuccess ||
        hipMalloc(&P->net_scratch, prepare_scratch_bytes(D.height, D.width, kCnnFactor)) != hipSuccess)
      return fail(PIV_E_NOMEM, "device allocation failed");
  }
  piv_pass_desc g{};
  g.search_size = 2 * kCnnFactor;  // node (i, j) at (2f j + f, 2f i + f)
  g.spacing = 2 * kCnnFactor;
  g.field_m = ni.out_h;
  g.field_n = ni.out_w;
  P->net_grid = g;
  P->net = net;
  P->net_first = first_pass;
  P->max_reject = max_reject_fraction;
  return PIV_OK;
}

int piv_plan_set_mask(piv_plan* P, const uint8_t* mask, int height, int width) {
  if (!P) return fail(PIV_E_INVALID, "null plan");
  const piv_plan_desc& D = P->desc;
  HIP_TRY(hipSetDevice(P->device));
  HIP_TRY(hipStreamSynchronize(P->stream));
  if (mask && P->net) return fail(PIV_E_INVALID, "a frame mask cannot be combined with a learned predictor");
  if (mask && (height != D.height || width != D.width))
    return fail(PIV_E_INVALID, "mask is %d x %d, the frames are %d x %d", height, width, D.height, D.width);
  const int np = (int)P->passes.size();
  // Every check runs before the plan changes: a rejected mask leaves the
  // previous state (masked or not) untouched.
  std::vector<std::vector<uint8_t>> fms(mask ? np : 0);
  std::vector<int> n_masked(np, 0), slow_gcv(np, 0);
  int maxN = 0, maxM = 0;
  bool slow = false;
  for (int k = 0; mask && k < np; ++k) {
    const PassBuffers& B = P->passes[k];
    fms[k] = field_mask_of(B.d, mask, D.width);
    for (uint8_t x : fms[k]) n_masked[k] += x;
    if (k < np - 1 && B.d.smooth && n_masked[k] > 0) {
      if (B.d.field_m > smoothn_masked_max_dim() || B.d.field_n > smoothn_masked_max_dim())
        return fail(PIV_E_INVALID, "masked smoothing supports fields of up to %d x %d",
                    smoothn_masked_max_dim(), smoothn_masked_max_dim());
      // mean weight (gpu_smoothn.py:171, 367): the masked vectors are NaN
      slow_gcv[k] = !((double)(B.N - n_masked[k]) / B.N > 0.9);
      slow = slow || (slow_gcv[k] && !(B.d.smoothing_par > 0));
      maxN = std::max(maxN, B.N);
      maxM = std::max(maxM, B.d.field_m);
    }
  }
  // the captured sequence changes
  if (P->graph_exec) (void)hipGraphExecDestroy(P->graph_exec);
  if (P->graph) (void)hipGraphDestroy(P->graph);
  P->graph_exec = nullptr;
  P->graph = nullptr;
  P->runs = 0;
  auto unmask = [P]() {
    for (auto& B : P->passes) {
      if (B.fmask) (void)hipFree(B.fmask);
      B.fmask = nullptr;
      B.n_masked = 0;
      B.slow_gcv = 0;
    }
    free_masked_smooth(&P->msb);
    if (P->frame_mask) (void)hipFree(P->frame_mask);
    P->frame_mask = nullptr;
  };
  unmask();
  if (!mask) return PIV_OK;
  // allocation or copy failures from here on leave the plan unmasked
  auto undo = [&](int code) {
    unmask();
    return code;
  };
  const size_t npx = (size_t)D.height * D.width;
  for (int k = 0; k < np; ++k) {
    PassBuffers& B = P->passes[k];
    if (hipMalloc(&B.fmask, fms[k].size()) != hipSuccess) {
      B.fmask = nullptr;
      return undo(fail(PIV_E_NOMEM, "device allocation failed"));
    }
    if (hipMemcpy(B.fmask, fms[k].data(), fms[k].size(), hipMemcpyHostToDevice) != hipSuccess)
      return undo(fail(PIV_E_HIP, "hipMemcpy failed"));
  }
  if (hipMalloc(&P->frame_mask, npx) != hipSuccess) {
    P->frame_mask = nullptr;
    return undo(fail(PIV_E_NOMEM, "device allocation failed"));
  }
  std::vector<uint8_t> fmk(npx);
  for (size_t q = 0; q < npx; ++q) fmk[q] = mask[q] ? 1 : 0;
  if (hipMemcpy(P->frame_mask, fmk.data(), npx, hipMemcpyHostToDevice) != hipSuccess)
    return undo(fail(PIV_E_HIP, "hipMemcpy failed"));
  if (maxN > 0) {
    int r = alloc_masked_smooth(maxN, maxM, slow, &P->msb, P->esize);
    if (r) return undo(r);
  }
  if (!P->smooth_iters) {
    if (hipMalloc(&P->smooth_iters, np * sizeof(int32_t)) != hipSuccess) {
      P->smooth_iters = nullptr;
      return undo(fail(PIV_E_NOMEM, "device allocation failed"));
    }
    if (hipMemset(P->smooth_iters, 0, np * sizeof(int32_t)) != hipSuccess)
      return undo(fail(PIV_E_HIP, "hipMemset failed"));
  }
  // the masked state becomes visible only once every buffer is in place
  for (int k = 0; k < np; ++k) {
    P->passes[k].n_masked = n_masked[k];
    P->passes[k].slow_gcv = slow_gcv[k];
  }
  return PIV_OK;
}

int piv_plan_last_smoothing_iterations(const piv_plan* P, int pass, int32_t* iterations) {
  if (!P || !iterations || pass < 0 || pass >= (int)P->passes.size()) return fail(PIV_E_INVALID, "bad argument");
  *iterations = 0;
  if (!P->smooth_iters) return PIV_OK;
  HIP_TRY(hipSetDevice(P->device));
  HIP_TRY(hipMemcpy(iterations, P->smooth_iters + pass, sizeof(int32_t), hipMemcpyDeviceToHost));
  return PIV_OK;
}

int piv_plan_last_s2n(piv_plan* P, void* s2n_out) {
  if (!P || !s2n_out) return fail(PIV_E_INVALID, "null argument");
  if (P->runs == 0 && P->lruns == 0) return fail(PIV_E_INVALID, "no pair has been processed yet");
  HIP_TRY(hipSetDevice(P->device));
  hipStream_t s = P->stream;
  HIP_TRY(hipStreamSynchronize(s));
  const int k = (int)P->passes.size() - 1;
  PassBuffers& B = P->passes[k];
  const piv_plan_desc& D = P->desc;
  // the last pass read the raw (masked) frames when it is the only one, the
  // deformed frames otherwise; both are still in the plan's buffers
  const bool raw = k == 0;
  const void* fa = raw ? src_a(P) : P->def_a;
  const void* fb = raw ? src_b(P) : P->def_b;
  const int dtype = raw ? D.in_dtype : (P->fp64 ? kF64 : kF32);
  void *tu = nullptr, *tv = nullptr, *ts = nullptr;
  const size_t bytes = (size_t)B.N * P->esize;
  int r = PIV_OK;
  if (hipMalloc(&tu, bytes) != hipSuccess || hipMalloc(&tv, bytes) != hipSuccess || hipMalloc(&ts, bytes) != hipSuccess)
    r = fail(PIV_E_NOMEM, "device allocation failed");
  hipError_t e = hipSuccess;
  if (!r) {
    if (P->fp64) {
      CorrArgs64 a = corr_args64(P, fa, fb, dtype, B.d);
      if (!raw) {
        a.up = as<double>(B.up);
        a.vp = as<double>(B.vp);
      }
      a.u = as<double>(tu);
      a.v = as<double>(tv);
      a.s2n = as<double>(ts);
      e = launch_correlate64(a, s);
      if (e == hipSuccess && B.fmask)
        e = launch_apply_field_mask(as<double>(tu), as<double>(tv), as<double>(ts), B.fmask, B.N, s);
    } else {
      CorrArgs a = corr_args(fa, fb, dtype, D.height, D.width, B.d);
      a.scratch = P->corr_scratch;
      if (!raw) {
        a.up = B.up;
        a.vp = B.vp;
      }
      a.u = as<float>(tu);
      a.v = as<float>(tv);
      a.s2n = as<float>(ts);
      e = launch_correlate(a, s);
      if (e == hipSuccess && B.fmask)
        e = launch_apply_field_mask(as<float>(tu), as<float>(tv), as<float>(ts), B.fmask, B.N, s);
    }
    if (e == hipSuccess) e = hipMemcpyAsync(s2n_out, ts, bytes, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
  }
  for (void* p : {tu, tv, ts})
    if (p) (void)hipFree(p);
  if (r) return r;
  if (e != hipSuccess) return fail(PIV_E_HIP, "last_s2n: %s [%s]", hipGetErrorString(e), launch_detail());
  return PIV_OK;
}

int piv_plan_wait_event(piv_plan* P, void* event) {
  if (!P || !event) return fail(PIV_E_INVALID, "null argument");
  P->wait_next = (hipEvent_t)event;
  return PIV_OK;
}

int piv_plan_last_fallback(const piv_plan* P, int32_t* fell_back, int32_t* rejected) {
  if (!P) return fail(PIV_E_INVALID, "null plan");
  if (fell_back) *fell_back = P->last_fallback;
  if (rejected) *rejected = P->last_rejected;
  return PIV_OK;
}

int piv_plan_last_smoothing(const piv_plan* P, int pass, double* log10_s) {
  if (!P || !log10_s || pass < 0 || pass >= (int)P->passes.size()) return fail(PIV_E_INVALID, "bad argument");
  HIP_TRY(hipSetDevice(P->device));
  HIP_TRY(hipMemcpy(log10_s, P->p_out + pass, sizeof(double), hipMemcpyDeviceToHost));
  return PIV_OK;
}

int piv_plan_last_counts(const piv_plan* P, int pass, int32_t* n_before, int32_t* n_after) {
  if (!P || pass < 0 || pass >= (int)P->passes.size()) return fail(PIV_E_INVALID, "bad argument");
  HIP_TRY(hipSetDevice(P->device));
  int32_t c[2];
  HIP_TRY(hipMemcpy(c, P->counters + 2 * pass, sizeof(c), hipMemcpyDeviceToHost));
  if (n_before) *n_before = c[0];
  if (n_after) *n_after = c[1];
  return PIV_OK;
}

int piv_stage_correlate(const void* frame_a, const void* frame_b, int in_dtype, int height, int width,
                        const piv_pass_desc* pass, const float* up, const float* vp, float* u, float* v,
                        int32_t* i_peak, int32_t* j_peak, float* planes, float* s2n, void* stream) {
  if (!frame_a || !frame_b || !pass || !u || !v) return fail(PIV_E_INVALID, "null argument");
  if (in_dtype < PIV_U8 || in_dtype > PIV_F32) return fail(PIV_E_INVALID, "in_dtype");
  int r = check_pass(*pass, height, width);
  if (r) return r;
  CorrArgs a = corr_args(frame_a, frame_b, in_dtype, height, width, *pass);
  a.up = up;
  a.vp = vp;
  a.u = u;
  a.v = v;
  a.i_peak = i_peak;
  a.j_peak = j_peak;
  a.planes = planes;
  a.s2n = s2n;
  const size_t sb = correlate_big_scratch_bytes(pass->fft_size, 4, pass->field_m * pass->field_n);
  if (!sb) {
    HIP_TRY(launch_correlate(a, (hipStream_t)stream));
    return PIV_OK;
  }
  // FFT sizes above 128: plane slabs for this call, released after it
  hipStream_t s = (hipStream_t)stream;
  if (hipMalloc(&a.scratch, sb) != hipSuccess) return fail(PIV_E_NOMEM, "device allocation failed");
  const hipError_t e = launch_correlate(a, s);
  const hipError_t e2 = hipStreamSynchronize(s);
  (void)hipFree(a.scratch);
  if (e != hipSuccess) return fail(PIV_E_HIP, "correlate: %s [%s]", hipGetErrorString(e), launch_detail());
  if (e2 != hipSuccess) return fail(PIV_E_HIP, "correlate: %s", hipGetErrorString(e2));
  return PIV_OK;
}

int piv_stage_deform(const void* frame_a, const void* frame_b, int in_dtype, int height, int width,
                     const piv_pass_desc* pass, const float* up, const float* vp, float* out_a, float* out_b,
                     void* stream) {
  if (!frame_a || !frame_b || !pass || !up || !vp || !out_a || !out_b) return fail(PIV_E_INVALID, "null argument");
  int r = check_pass(*pass, height, width);
  if (r) return r;
  if (pass->field_m < 2 || pass->field_n < 2) return fail(PIV_E_INVALID, "deformation needs a 2 x 2 grid");
  HIP_TRY(launch_deform(deform_args(frame_a, frame_b, in_dtype, height, width, *pass, up, vp, out_a, out_b),
                        (hipStream_t)stream));
  return PIV_OK;
}

int piv_stage_grid_interp(const piv_pass_desc* from, const piv_pass_desc* to, const float* u_in, const float* v_in,
                          float* u_out, float* v_out, void* stream) {
  if (!from || !to || !u_in || !v_in || !u_out || !v_out) return fail(PIV_E_INVALID, "null argument");
  if (from->field_m < 2 || from->field_n < 2) return fail(PIV_E_INVALID, "source grid must be at least 2 x 2");
  HIP_TRY(launch_grid_interp(interp_args(*from, *to, u_in, v_in, u_out, v_out), (hipStream_t)stream));
  return PIV_OK;
}

int piv_stage_median_validate(const float* u, const float* v, int m, int n, float median_tol, uint8_t* flags,
                              int32_t* count, void* stream) {
  if (!u || !v || !flags || !count || m < 1 || n < 1) return fail(PIV_E_INVALID, "bad argument");
  HIP_TRY(launch_median_validate(u, v, m, n, median_tol, flags, count, (hipStream_t)stream));
  return PIV_OK;
}

int piv_stage_validate(const float* u, const float* v, int m, int n, const float* s2n, const piv_pass_desc* pass,
                       const uint8_t* only, uint8_t* flags, int32_t* count, void* stream) {
  if (!u || !v || !pass || !flags || !count || m < 1 || n < 1) return fail(PIV_E_INVALID, "bad argument");
  int r = check_tests(*pass, m, n);
  if (r) return r;
  double* stats = nullptr;
  hipStream_t s = (hipStream_t)stream;
  if (!std::isnan(pass->rms_tol)) HIP_TRY(hipMalloc((void**)&stats, 4 * sizeof(double)));
  r = run_validate(*pass, u, v, m, n, s2n, stats, only, nullptr, flags, count, s);
  if (stats) {
    (void)hipStreamSynchronize(s);
    (void)hipFree(stats);
  }
  return r;
}

int piv_stage_replace(float* u, float* v, uint8_t* flags, int m, int n, const piv_pass_desc* pass, const float* s2n,
                      void* stream) {
  if (!u || !v || !flags || !pass || m < 1 || n < 1) return fail(PIV_E_INVALID, "bad argument");
  int r = check_tests(*pass, m, n);
  if (r) return r;
  hipStream_t s = (hipStream_t)stream;
  const size_t N = (size_t)m * n;
  LoopBufs L{};
  int32_t* count0 = nullptr;
  std::vector<void*> owned;
  auto grab = [&](void** p, size_t bytes) -> bool {
    if (hipMalloc(p, bytes) != hipSuccess) return false;
    owned.push_back(*p);
    return true;
  };
  bool ok = grab((void**)&L.u0, N * 4) && grab((void**)&L.v0, N * 4) && grab((void**)&L.tu, N * 4) &&
            grab((void**)&L.tv, N * 4) && grab((void**)&L.wf[0], N) && grab((void**)&L.wf[1], N) &&
            grab((void**)&L.nanf, N) && grab((void**)&L.it_counts, PIV_MAX_REPLACING_ITERS * 4) &&
            grab((void**)&L.stats, 32) && grab((void**)&L.rw.slot, N * 4) && grab((void**)&L.rw.list, N * 4) &&
            grab((void**)&L.rw.work, 8 * N * 8) && grab((void**)&count0, 4);
  auto cleanup = [&] {
    (void)hipStreamSynchronize(s);
    for (void* p : owned) (void)hipFree(p);
  };
  if (!ok) {
    cleanup();
    return fail(PIV_E_NOMEM, "device allocation failed");
  }
  r = enqueue_stage_replace(u, v, flags, m, n, *pass, s2n, L, count0, s);
  cleanup();
  return r;
}

int piv_stage_replace_noalloc(float* u, float* v, uint8_t* flags, int m, int n, const piv_pass_desc* pass,
                              const float* s2n, void* stream) {
  if (!u || !v || !flags || !pass || m < 1 || n < 1) return fail(PIV_E_INVALID, "bad argument");
  if ((size_t)m * n > (size_t)PIV_NOALLOC_MAX_VECTORS)
    return fail(PIV_E_INVALID, "piv_stage_replace_noalloc: %d x %d vectors exceed its fixed workspace of %d vectors",
                m, n, PIV_NOALLOC_MAX_VECTORS);
  int r = check_tests(*pass, m, n);
  if (r) return r;
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  unsigned char* base = nullptr;
  r = noalloc_scratch(dev, &base);
  if (r) return r;
  LoopBufs L{};
  int32_t* count0 = nullptr;
  noalloc_layout(base, (size_t)m * n, &L, &count0);
  return enqueue_stage_replace(u, v, flags, m, n, *pass, s2n, L, count0, (hipStream_t)stream);
}

int piv_stage_spring_replace(float* u, float* v, const uint8_t* flags, const int32_t* count, int m, int n,
                             void* stream) {
  if (!u || !v || !flags || !count || m < 1 || n < 1) return fail(PIV_E_INVALID, "bad argument");
  const size_t N = (size_t)m * n;
  ReplaceWork w{};
  hipStream_t s = (hipStream_t)stream;
  HIP_TRY(hipMalloc((void**)&w.slot, N * 4));
  HIP_TRY(hipMalloc((void**)&w.list, N * 4));
  HIP_TRY(hipMalloc((void**)&w.work, 8 * N * 8));
  hipError_t e = launch_spring_replace(u, v, flags, count, m, n, 1, w, s);
  hipError_t e2 = hipStreamSynchronize(s);
  (void)hipFree(w.slot);
  (void)hipFree(w.list);
  (void)hipFree(w.work);
  if (e != hipSuccess) return fail(PIV_E_HIP, "spring replace: %s", hipGetErrorString(e));
  if (e2 != hipSuccess) return fail(PIV_E_HIP, "spring replace: %s", hipGetErrorString(e2));
  return PIV_OK;
}

int piv_stage_smoothn(float* u, float* v, int m, int n, double smoothing_par, double* log10_s_out, void* stream) {
  if (!u || !v || m < 1 || n < 1) return fail(PIV_E_INVALID, "bad argument");
  hipStream_t s = (hipStream_t)stream;
  SmoothTables tab;
  int r = make_tables(m, n, &tab);
  if (r) {
    free_tables(&tab);
    return r;
  }
  const size_t N = (size_t)m * n;
  float *t0 = nullptr, *t1 = nullptr;
  void* state = nullptr;
  HIP_TRY(hipMalloc(&t0, 2 * N * 4));
  HIP_TRY(hipMalloc(&t1, 2 * N * 4));
  HIP_TRY(hipMalloc(&state, smoothn_state_bytes()));
  HIP_TRY(hipMemset(state, 0, smoothn_state_bytes()));
  double pmin, pmax;
  p_bounds(m, n, &pmin, &pmax);
  SmoothnWork w{(const float*)tab.cos_m, (const float*)tab.cos_n, tab.lambda, t0, t1, log10_s_out, state};
  const int autos = !(smoothing_par > 0);
  hipError_t e = launch_smoothn(u, v, m, n, pmin, pmax, autos ? 0.0 : log10(smoothing_par), autos, w, s);
  hipError_t e2 = hipStreamSynchronize(s);
  (void)hipFree(t0);
  (void)hipFree(t1);
  (void)hipFree(state);
  free_tables(&tab);
  if (e != hipSuccess) return fail(PIV_E_HIP, "smoothn: %s", hipGetErrorString(e));
  if (e2 != hipSuccess) return fail(PIV_E_HIP, "smoothn: %s", hipGetErrorString(e2));
  return PIV_OK;
}

int piv_stage_correlate64(const void* frame_a, const void* frame_b, int in_dtype, int height, int width,
                          const piv_pass_desc* pass, const double* up, const double* vp, double* u, double* v,
                          int32_t* i_peak, int32_t* j_peak, double* planes, double* s2n, void* stream) {
  if (!frame_a || !frame_b || !pass || !u || !v) return fail(PIV_E_INVALID, "null argument");
  if (in_dtype < PIV_U8 || in_dtype > PIV_F64) return fail(PIV_E_INVALID, "in_dtype");
  int r = check_pass(*pass, height, width);
  if (r) return r;
  hipStream_t s = (hipStream_t)stream;
  double* tw = nullptr;
  double2* scratch = nullptr;
  r = upload((void**)&tw, fft_twiddles(pass->fft_size));
  const size_t sb = correlate64_scratch_bytes(pass->fft_size);
  if (!r && sb && hipMalloc((void**)&scratch, sb) != hipSuccess) r = fail(PIV_E_NOMEM, "device allocation failed");
  hipError_t e = hipSuccess, e2 = hipSuccess;
  if (!r) {
    CorrArgs64 a = corr_args64_for(frame_a, frame_b, in_dtype, height, width, *pass, tw, scratch);
    a.up = up;
    a.vp = vp;
    a.u = u;
    a.v = v;
    a.i_peak = i_peak;
    a.j_peak = j_peak;
    a.planes = planes;
    a.s2n = s2n;
    e = launch_correlate64(a, s);
    e2 = hipStreamSynchronize(s);
  }
  if (tw) (void)hipFree(tw);
  if (scratch) (void)hipFree(scratch);
  if (r) return r;
  if (e != hipSuccess) return fail(PIV_E_HIP, "correlate64: %s [%s]", hipGetErrorString(e), launch_detail());
  if (e2 != hipSuccess) return fail(PIV_E_HIP, "correlate64: %s", hipGetErrorString(e2));
  return PIV_OK;
}

int piv_stage_smoothn_masked(float* u, float* v, const uint8_t* mask, int m, int n, double smoothing_par,
                             double* log10_s_out, int32_t* iterations_out, void* stream) {
  if (!u || !v || m < 1 || n < 1) return fail(PIV_E_INVALID, "bad argument");
  if (m > smoothn_masked_max_dim() || n > smoothn_masked_max_dim())
    return fail(PIV_E_INVALID, "masked smoothing supports fields of up to %d x %d", smoothn_masked_max_dim(),
                smoothn_masked_max_dim());
  hipStream_t s = (hipStream_t)stream;
  const int N = m * n;
  // the branch of the GCV score follows the mean weight, from the mask and
  // the non-finite vectors (gpu_smoothn.py:171, 259-269, 367)
  std::vector<float> hu(N), hv(N);
  std::vector<uint8_t> hm(N, 0);
  HIP_TRY(hipMemcpyAsync(hu.data(), u, (size_t)N * 4, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(hv.data(), v, (size_t)N * 4, hipMemcpyDeviceToHost, s));
  if (mask) HIP_TRY(hipMemcpyAsync(hm.data(), mask, N, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  int n_w = 0;
  for (int k = 0; k < N; ++k) n_w += (std::isfinite(hu[k]) && std::isfinite(hv[k]) && !hm[k]) ? 1 : 0;
  const int slow = !((double)n_w / N > 0.9);
  SmoothTables tab;
  int r = make_tables(m, n, &tab);
  MaskedSmoothBufs b;
  float *t0 = nullptr, *t1 = nullptr;
  if (!r) r = alloc_masked_smooth(N, m, slow != 0, &b);
  if (!r && (hipMalloc(&t0, 2ull * N * 4) != hipSuccess || hipMalloc(&t1, 2ull * N * 4) != hipSuccess))
    r = fail(PIV_E_NOMEM, "device allocation failed");
  hipError_t e = hipSuccess, e2 = hipSuccess;
  if (!r) {
    double pmin, pmax;
    p_bounds(m, n, &pmin, &pmax);
    const int autos = !(smoothing_par > 0);
    const SmoothnMaskedWork w = masked_work<float>(tab, mask, b, t0, t1, log10_s_out, iterations_out, slow);
    if (n_w == N) {
      // nothing masked or missing: the unweighted single pass
      void* state = b.state;
      SmoothnWork uw{(const float*)tab.cos_m, (const float*)tab.cos_n, tab.lambda, t0, t1, log10_s_out, state};
      e = launch_smoothn(u, v, m, n, pmin, pmax, autos ? 0.0 : log10(smoothing_par), autos, uw, s);
      if (e == hipSuccess && iterations_out) e = hipMemsetAsync(iterations_out, 0, 4, s);
    } else {
      e = launch_smoothn_masked(u, v, m, n, pmin, pmax, autos ? 0.0 : log10(smoothing_par), autos, w, s);
    }
    e2 = hipStreamSynchronize(s);
  }
  if (t0) (void)hipFree(t0);
  if (t1) (void)hipFree(t1);
  free_masked_smooth(&b);
  free_tables(&tab);
  if (r) return r;
  if (e != hipSuccess) return fail(PIV_E_HIP, "smoothn: %s", hipGetErrorString(e));
  if (e2 != hipSuccess) return fail(PIV_E_HIP, "smoothn: %s", hipGetErrorString(e2));
  return PIV_OK;
}

int piv_stage_correlate_dft(const float* win_a, const float* win_b, int n_windows, int window_size, int fft_size,
                            int precision, int subpixel, float* u, float* v, int32_t* i_peak, int32_t* j_peak,
                            float* planes, void* stream) {
  if (!win_a || !win_b || !u || !v || n_windows < 0) return fail(PIV_E_INVALID, "bad argument");
  const int W = window_size, F = fft_size;
  if (W < 16 || W % 16 || !is_pow2(F) || F < W || (F - W) % 2 || F > 256)
    return fail(PIV_E_INVALID, "window_size must be a multiple of 16 and fft_size a power of two >= it");
  if (precision < 0 || precision > 3 || subpixel < 0 || subpixel > 2) return fail(PIV_E_INVALID, "mode");
  if (n_windows == 0) return PIV_OK;
  hipStream_t s = (hipStream_t)stream;
  const int pad = (F - W) / 2;
  std::vector<float> er((size_t)F * W), ei((size_t)F * W), gr((size_t)F * F), gi((size_t)F * F);
  for (int k = 0; k < F; ++k) {
    for (int n = 0; n < W; ++n) {
      const double a = -2.0 * M_PI * (double)((long)k * (n + pad) % F) / F;
      er[(size_t)k * W + n] = (float)cos(a);
      ei[(size_t)k * W + n] = (float)sin(a);
    }
    for (int x = 0; x < F; ++x) {
      const double a = 2.0 * M_PI * (double)((long)k * x % F) / F;
      gr[(size_t)x * F + k] = (float)cos(a);
      gi[(size_t)x * F + k] = (float)sin(a);
    }
  }
  // process in chunks that keep the workspace near 1 GiB
  const size_t per = dft_workspace_bytes(1, W, F);
  const int chunk = (int)std::max<size_t>(1, std::min<size_t>((size_t)n_windows, ((size_t)1 << 30) / per));
  float *d_er = nullptr, *d_ei = nullptr, *d_gr = nullptr, *d_gi = nullptr, *work = nullptr;
  auto cleanup = [&] {
    for (void* p : {(void*)d_er, (void*)d_ei, (void*)d_gr, (void*)d_gi, (void*)work})
      if (p) (void)hipFree(p);
  };
  if (hipMalloc(&d_er, er.size() * 4) != hipSuccess || hipMalloc(&d_ei, ei.size() * 4) != hipSuccess ||
      hipMalloc(&d_gr, gr.size() * 4) != hipSuccess || hipMalloc(&d_gi, gi.size() * 4) != hipSuccess ||
      hipMalloc((void**)&work, per * chunk) != hipSuccess) {
    cleanup();
    return fail(PIV_E_NOMEM, "DFT workspace allocation failed");
  }
  hipError_t e = hipMemcpy(d_er, er.data(), er.size() * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d_ei, ei.data(), ei.size() * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d_gr, gr.data(), gr.size() * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d_gi, gi.data(), gi.size() * 4, hipMemcpyHostToDevice);
  for (int b0 = 0; b0 < n_windows && e == hipSuccess; b0 += chunk) {
    const int nb = std::min(chunk, n_windows - b0);
    DftArgs d{};
    d.win_a = win_a + (size_t)b0 * W * W;
    d.win_b = win_b + (size_t)b0 * W * W;
    d.n = nb;
    d.ws = W;
    d.fsize = F;
    d.precision = precision;
    d.subpixel = subpixel;
    d.er = d_er;
    d.ei = d_ei;
    d.gr = d_gr;
    d.gi = d_gi;
    d.workspace = work;
    d.u = u + b0;
    d.v = v + b0;
    d.i_peak = i_peak ? i_peak + b0 : nullptr;
    d.j_peak = j_peak ? j_peak + b0 : nullptr;
    d.planes = planes ? planes + (size_t)b0 * F * F : nullptr;
    e = launch_dft_correlate(d, s);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  cleanup();
  if (e != hipSuccess) return fail(PIV_E_HIP, "DFT correlation: %s", hipGetErrorString(e));
  return PIV_OK;
}

int piv_set_device(int device) {
  HIP_TRY(hipSetDevice(device));
  return PIV_OK;
}
int piv_malloc(void** ptr, size_t bytes) {
  if (!ptr) return fail(PIV_E_INVALID, "null argument");
  *ptr = nullptr;
  hipError_t e = hipMalloc(ptr, bytes ? bytes : 16);
  if (e != hipSuccess) return fail(PIV_E_NOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
  return PIV_OK;
}
int piv_free(void* ptr) {
  if (ptr) HIP_TRY(hipFree(ptr));
  return PIV_OK;
}
int piv_memcpy(void* dst, const void* src, size_t bytes, int kind) {
  if ((!dst || !src) && bytes) return fail(PIV_E_INVALID, "null argument");
  const hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice
                                    : (kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice);
  if (bytes) HIP_TRY(hipMemcpy(dst, src, bytes, k));
  return PIV_OK;
}
int piv_memset(void* ptr, int value, size_t bytes) {
  if (!ptr && bytes) return fail(PIV_E_INVALID, "null argument");
  if (bytes) HIP_TRY(hipMemset(ptr, value, bytes));
  return PIV_OK;
}
int piv_synchronize(void) {
  HIP_TRY(hipDeviceSynchronize());
  return PIV_OK;
}
int piv_stream_create(void** stream) {
  if (!stream) return fail(PIV_E_INVALID, "null argument");
  hipStream_t s = nullptr;
  HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *stream = (void*)s;
  return PIV_OK;
}
int piv_stream_destroy(void* stream) {
  if (stream) HIP_TRY(hipStreamDestroy((hipStream_t)stream));
  return PIV_OK;
}
int piv_stream_synchronize(void* stream) {
  HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
  return PIV_OK;
}
int piv_event_create(void** event) {
  if (!event) return fail(PIV_E_INVALID, "null argument");
  hipEvent_t e = nullptr;
  HIP_TRY(hipEventCreate(&e));
  *event = (void*)e;
  return PIV_OK;
}
int piv_event_destroy(void* event) {
  if (event) HIP_TRY(hipEventDestroy((hipEvent_t)event));
  return PIV_OK;
}
int piv_event_record(void* event, void* stream) {
  if (!event) return fail(PIV_E_INVALID, "null argument");
  HIP_TRY(hipEventRecord((hipEvent_t)event, (hipStream_t)stream));
  return PIV_OK;
}
int piv_event_elapsed_ms(void* start, void* stop, float* ms) {
  if (!start || !stop || !ms) return fail(PIV_E_INVALID, "null argument");
  HIP_TRY(hipEventSynchronize((hipEvent_t)stop));
  HIP_TRY(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
  return PIV_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batches of pairs with their own frame shapes and schedules.  Pass level L
// of every pair runs as: per-pair predictor interpolation, ONE deformation
// launch over all pairs' frames, ONE correlation launch per (FFT size, input
// dtype) over all pairs' windows, then the per-pair field stages
// (validation, replacement, smoothn).  The whole loop is one hipGraph after
// the first (eager) run.  Launch shapes per level (XCD-banded or linear
// block order for the correlation, block width of the deformation) come from
// piv_batch_tune, which times the candidates on the batch's own data.
// ---------------------------------------------------------------------------

namespace {

struct BatchGroup {  // one batched correlation launch
  int fsize = 0, in_dtype = 0, total_blocks = 0, max_ss = 0, xcd = 1;
  int bw = 1;  // line kernel: waves per workgroup (tuned)
  bool line = false;  // one-lane-per-line kernel (the one a single-pair run of these fields uses)
  bool split = false;  // residue-split kernel (likewise)
  bool half = false;  // search windows <= F/2 (the two-lane kernel's variant for them)
  std::vector<int> plans;
  CorrArgs* d_fields = nullptr;
  int* d_prefix = nullptr;
  // line groups: block prefixes for workgroups of 1, 2 and 4 waves
  int* d_prefix_bw[3] = {nullptr, nullptr, nullptr};
  int total_blocks_bw[3] = {0, 0, 0};
};
constexpr int kLineWidths[3] = {1, 2, 4};
int line_width_index(int bw) { return bw == 4 ? 2 : (bw == 2 ? 1 : 0); }

// One launch of a recorded field-stage kernel over a group of pairs whose
// recorded sequences match (piv_record.h).
struct BatchedStep {
  LaunchKind kind = LaunchKind::kFill;
  int n = 0, total_blocks = 0;
  void* d_args = nullptr;    // n argument structs
  int4* d_grids = nullptr;   // n (gx, gy, gz, first block)
};

struct BatchLevel {
  std::vector<int> plans;  // pairs with a pass at this level
  std::vector<BatchGroup> groups;
  DeformArgs* d_deform = nullptr;  // pairs with L > 0
  int* d_rows = nullptr;           // first row group of each pair's frames
  int n_deform = 0, total_rows = 0, max_wd = 0, deform_threads = 256;  // total_rows: row groups
  std::vector<BatchedStep> pre;   // predictor interpolation (L > 0)
  std::vector<BatchedStep> post;  // validation, replacement, smoothing
};

}  // namespace

struct piv_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<piv_plan*> plans;
  std::vector<BatchLevel> levels;
  bool frames_set = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  int runs = 0;
  std::vector<BatchedStep> start;   // counter resets
  std::vector<BatchedStep> finish;  // output scaling
};

namespace {

void free_steps(std::vector<BatchedStep>& steps) {
  for (auto& st : steps) {
    if (st.d_args) (void)hipFree(st.d_args);
    if (st.d_grids) (void)hipFree(st.d_grids);
  }
  steps.clear();
}

void batch_free_tables(piv_batch* b) {
  for (auto& L : b->levels) {
    for (auto& g : L.groups) {
      if (g.d_fields) (void)hipFree(g.d_fields);
      for (int* p : g.d_prefix_bw)
        if (p && p != g.d_prefix) (void)hipFree(p);
      if (g.d_prefix) (void)hipFree(g.d_prefix);
    }
    if (L.d_deform) (void)hipFree(L.d_deform);
    if (L.d_rows) (void)hipFree(L.d_rows);
    free_steps(L.pre);
    free_steps(L.post);
  }
  free_steps(b->start);
  free_steps(b->finish);
}

// The launches fn() makes, recorded instead of launched.
template <typename Fn>
int record_sequence(std::vector<RecordedLaunch>* seq, Fn&& fn) {
  g_launch_record = seq;
  const int r = fn();
  g_launch_record = nullptr;
  return r;
}

// Batched steps from the recorded sequences of several pairs: pairs whose
// sequences have the same kinds in the same order form a group, and step j
// of a group is one launch over the group's launch j.
int build_steps(const std::vector<std::vector<RecordedLaunch>>& seqs, std::vector<BatchedStep>& out) {
  std::vector<bool> taken(seqs.size(), false);
  for (size_t i = 0; i < seqs.size(); ++i) {
    if (taken[i] || seqs[i].empty()) continue;
    std::vector<size_t> group;
    for (size_t q = i; q < seqs.size(); ++q) {
      if (taken[q] || seqs[q].size() != seqs[i].size()) continue;
      bool same = true;
      for (size_t j = 0; j < seqs[i].size() && same; ++j)
        same = seqs[q][j].kind == seqs[i][j].kind && seqs[q][j].block == seqs[i][j].block &&
               seqs[q][j].args.size() == seqs[i][j].args.size();
      if (same) {
        group.push_back(q);
        taken[q] = true;
      }
    }
    for (size_t j = 0; j < seqs[i].size(); ++j) {
      BatchedStep st;
      st.kind = seqs[i][j].kind;
      const size_t asz = seqs[i][j].args.size();
      std::vector<unsigned char> args;
      std::vector<int4> grids;
      long long blocks = 0;
      for (size_t q : group) {
        const RecordedLaunch& r = seqs[q][j];
        args.insert(args.end(), r.args.begin(), r.args.end());
        grids.push_back(make_int4(r.gx, r.gy, r.gz, (int)blocks));
        blocks += (long long)r.gx * r.gy * r.gz;
      }
      if (blocks > INT_MAX) return fail(PIV_E_INVALID, "batch: too many blocks in one field-stage launch");
      st.n = (int)group.size();
      st.total_blocks = (int)blocks;
      if (hipMalloc(&st.d_args, args.size()) != hipSuccess ||
          hipMalloc((void**)&st.d_grids, grids.size() * sizeof(int4)) != hipSuccess) {
        out.push_back(st);  // freed with the tables
        return fail(PIV_E_NOMEM, "batch: table allocation failed");
      }
      out.push_back(st);
      HIP_TRY(hipMemcpy(st.d_args, args.data(), st.n * asz, hipMemcpyHostToDevice));
      HIP_TRY(hipMemcpy(st.d_grids, grids.data(), grids.size() * sizeof(int4), hipMemcpyHostToDevice));
    }
  }
  return PIV_OK;
}

int run_steps(const std::vector<BatchedStep>& steps, hipStream_t s) {
  for (const BatchedStep& st : steps) {
    const hipError_t e = st.kind >= LaunchKind::kGemm
                             ? launch_smoothn_batch(st.kind, st.d_args, st.d_grids, st.n, st.total_blocks, s)
                             : launch_field_batch(st.kind, st.d_args, st.d_grids, st.n, st.total_blocks, s);
    if (e != hipSuccess) return fail(PIV_E_HIP, "batch: field-stage launch: %s", hipGetErrorString(e));
  }
  return PIV_OK;
}

// Records every pair's field-stage launches (start, per level before and
// after the correlation, finish) and builds the batched steps.
int batch_build_steps(piv_batch* b) {
  const size_t np = b->plans.size();
  std::vector<std::vector<RecordedLaunch>> seqs(np);
  for (size_t i = 0; i < np; ++i) {
    piv_plan* P = b->plans[i];
    int r = record_sequence(&seqs[i], [&] {
      HIP_TRY(launch_fill_i32(P->counters, 0, 2 * (int)P->passes.size(), nullptr));
      HIP_TRY(launch_fill_i32(P->lb.it_counts, 0, PIV_MAX_REPLACING_ITERS * (int)P->passes.size(), nullptr));
      return (int)PIV_OK;
    });
    if (r) return r;
  }
  int r = build_steps(seqs, b->start);
  if (r) return r;
  for (size_t l = 0; l < b->levels.size(); ++l) {
    BatchLevel& L = b->levels[l];
    std::vector<std::vector<RecordedLaunch>> pre(L.plans.size()), post(L.plans.size());
    for (size_t q = 0; q < L.plans.size(); ++q) {
      piv_plan* P = b->plans[L.plans[q]];
      const bool last = l + 1 == P->passes.size();
      if (l > 0) r = record_sequence(&pre[q], [&] { return predictor_from_previous<float>(P, (int)l, nullptr); });
      if (!r) r = record_sequence(&post[q], [&] { return validate_and_smooth<float>(P, (int)l, last, nullptr); });
      if (r) return r;
    }
    r = build_steps(pre, L.pre);
    if (!r) r = build_steps(post, L.post);
    if (r) return r;
  }
  std::vector<std::vector<RecordedLaunch>> fin(np);
  for (size_t i = 0; i < np; ++i) {
    piv_plan* P = b->plans[i];
    if (P->desc.out_scale == 1.0f) continue;
    PassBuffers& Lp = P->passes.back();
    r = record_sequence(&fin[i], [&] {
      HIP_TRY(launch_scale(Lp.u, Lp.v, Lp.N, P->desc.out_scale, nullptr));
      return (int)PIV_OK;
    });
    if (r) return r;
  }
  return build_steps(fin, b->finish);
}

void batch_drop_graph(piv_batch* b) {
  if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->graph_exec = nullptr;
  b->graph = nullptr;
  b->runs = 0;
}

void batch_release(piv_batch* b) {
  if (!b) return;
  (void)hipSetDevice(b->device);
  batch_drop_graph(b);
  batch_free_tables(b);
  for (auto* P : b->plans) release(P);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

// Device tables of every level for the current frames (called once the
// frame pointers are known; the plans' own buffers never move).
int batch_build_tables(piv_batch* b) {
  batch_free_tables(b);
  for (size_t l = 0; l < b->levels.size(); ++l) {
    BatchLevel& L = b->levels[l];
    std::vector<int> keep_xcd, keep_bw;
    for (auto& g : L.groups) {
      keep_xcd.push_back(g.xcd);
      keep_bw.push_back(g.bw);
    }
    const int threads = L.deform_threads;
    L = BatchLevel{};
    L.deform_threads = threads;
    for (int p = 0; p < (int)b->plans.size(); ++p)
      if ((int)b->plans[p]->passes.size() > (int)l) L.plans.push_back(p);
    // correlation groups by (FFT size, input dtype, kernel)
    for (int p : L.plans) {
      piv_plan* P = b->plans[p];
      const PassBuffers& B = P->passes[l];
      const piv_plan_desc& D = P->desc;
      const int dtype = l == 0 ? D.in_dtype : (int)kF32;
      CorrArgs probe = l == 0 ? corr_args(src_a(P), src_b(P), D.in_dtype, D.height, D.width, B.d)
                              : corr_args(P->def_a, P->def_b, kF32, D.height, D.width, B.d);
      probe.s2n = B.s2n;
      const bool line = correlate_line_enabled() && correlate_line_applies(probe);
      const bool split = !line && correlate_split_enabled() && correlate_split_applies(probe);
      const bool half = 2 * B.d.search_size <= B.d.fft_size;
      BatchGroup* g = nullptr;
      for (auto& q : L.groups)
        if (q.fsize == B.d.fft_size && q.in_dtype == dtype && q.line == line && q.split == split && q.half == half)
          g = &q;
      if (!g) {
        L.groups.push_back(BatchGroup{});
        g = &L.groups.back();
        g->fsize = B.d.fft_size;
        g->in_dtype = dtype;
        g->line = line;
        g->split = split;
        g->half = half;
        g->xcd = L.groups.size() <= keep_xcd.size() ? keep_xcd[L.groups.size() - 1] : 1;
        g->bw = line && L.groups.size() <= keep_bw.size() ? keep_bw[L.groups.size() - 1] : 1;
      }
      g->plans.push_back(p);
    }
    for (auto& g : L.groups) {
      std::vector<CorrArgs> fields;
      std::vector<int> prefix;
      const int wpb = g.line    ? correlate_line_windows_per_block(g.fsize, g.bw)
                      : g.split ? correlate_split_windows_per_block(g.fsize)
                                : correlate_windows_per_block(g.fsize);
      g.total_blocks = 0;
      for (int p : g.plans) {
        piv_plan* P = b->plans[p];
        PassBuffers& B = P->passes[l];
        const piv_plan_desc& D = P->desc;
        CorrArgs a = l == 0 ? corr_args(src_a(P), src_b(P), D.in_dtype, D.height, D.width, B.d)
                            : corr_args(P->def_a, P->def_b, kF32, D.height, D.width, B.d);
        if (l > 0) {
          a.up = B.up;
          a.vp = B.vp;
        }
        a.u = B.u;
        a.v = B.v;
        a.s2n = B.s2n;
        fields.push_back(a);
        prefix.push_back(g.total_blocks);
        g.total_blocks += (B.N + wpb - 1) / wpb;
        g.max_ss = std::max(g.max_ss, B.d.search_size);
      }
      if (hipMalloc((void**)&g.d_fields, fields.size() * sizeof(CorrArgs)) != hipSuccess ||
          hipMalloc((void**)&g.d_prefix, prefix.size() * sizeof(int)) != hipSuccess)
        return fail(PIV_E_NOMEM, "batch: table allocation failed");
      HIP_TRY(hipMemcpy(g.d_fields, fields.data(), fields.size() * sizeof(CorrArgs), hipMemcpyHostToDevice));
      HIP_TRY(hipMemcpy(g.d_prefix, prefix.data(), prefix.size() * sizeof(int), hipMemcpyHostToDevice));
      if (g.line) {  // the prefixes of the other workgroup widths, for the tuner
        for (int wi = 0; wi < 3; ++wi) {
          if (kLineWidths[wi] == g.bw) {
            g.d_prefix_bw[wi] = g.d_prefix;
            g.total_blocks_bw[wi] = g.total_blocks;
            continue;
          }
          const int w = correlate_line_windows_per_block(g.fsize, kLineWidths[wi]);
          std::vector<int> pre;
          int tot = 0;
          for (int p : g.plans) {
            pre.push_back(tot);
            tot += (b->plans[p]->passes[l].N + w - 1) / w;
          }
          if (hipMalloc((void**)&g.d_prefix_bw[wi], pre.size() * sizeof(int)) != hipSuccess)
            return fail(PIV_E_NOMEM, "batch: table allocation failed");
          HIP_TRY(hipMemcpy(g.d_prefix_bw[wi], pre.data(), pre.size() * sizeof(int), hipMemcpyHostToDevice));
          g.total_blocks_bw[wi] = tot;
        }
      }
    }
    if (l > 0) {
      std::vector<DeformArgs> dfs;
      std::vector<int> rows;
      for (int p : L.plans) {
        piv_plan* P = b->plans[p];
        const PassBuffers& B = P->passes[l];
        const piv_plan_desc& D = P->desc;
        dfs.push_back(deform_args<float>(src_a(P), src_b(P), D.in_dtype, D.height, D.width, B.d, B.up, B.vp,
                                         as<float>(P->def_a), as<float>(P->def_b)));
        rows.push_back(L.total_rows);
        L.total_rows += (D.height + kDeformRows - 1) / kDeformRows;
        L.max_wd = std::max(L.max_wd, D.width);
      }
      L.n_deform = (int)dfs.size();
      if (L.n_deform > 0) {
        if (hipMalloc((void**)&L.d_deform, dfs.size() * sizeof(DeformArgs)) != hipSuccess ||
            hipMalloc((void**)&L.d_rows, rows.size() * sizeof(int)) != hipSuccess)
          return fail(PIV_E_NOMEM, "batch: table allocation failed");
        HIP_TRY(hipMemcpy(L.d_deform, dfs.data(), dfs.size() * sizeof(DeformArgs), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(L.d_rows, rows.data(), rows.size() * sizeof(int), hipMemcpyHostToDevice));
      }
    }
  }
  return batch_build_steps(b);
}

int batch_enqueue_corr(piv_batch* b, BatchGroup& g, hipStream_t s) {
  CorrBatch cb{g.d_fields, g.d_prefix, (int)g.plans.size(), g.xcd};
  if (g.line)
    HIP_TRY(launch_correlate_line_batch(cb, g.fsize, g.in_dtype, g.total_blocks, g.max_ss, g.bw, s));
  else if (g.split)
    HIP_TRY(launch_correlate_split_batch(cb, g.fsize, g.in_dtype, g.total_blocks, s));
  else
    HIP_TRY(launch_correlate_batch(cb, g.fsize, g.in_dtype, g.total_blocks, g.max_ss, g.half, s));
  return PIV_OK;
}

int batch_enqueue_deform(piv_batch* b, BatchLevel& L, hipStream_t s) {
  if (L.n_deform == 0) return PIV_OK;
  const int in_dtype = b->plans[L.plans[0]]->desc.in_dtype;
  HIP_TRY(launch_deform_batch(L.d_deform, L.d_rows, L.n_deform, L.total_rows, L.max_wd, in_dtype, L.deform_threads,
                              s));
  return PIV_OK;
}

// The batch's pass loop: per level one batched launch per field-stage
// kernel (the recorded steps), one deformation launch and one correlation
// launch per group.
int batch_enqueue(piv_batch* b) {
  hipStream_t s = b->stream;
  int r = run_steps(b->start, s);
  for (size_t l = 0; l < b->levels.size() && !r; ++l) {
    BatchLevel& L = b->levels[l];
    if (l > 0) {
      r = run_steps(L.pre, s);
      if (!r) r = batch_enqueue_deform(b, L, s);
    }
    for (auto& g : L.groups)
      if (!r) r = batch_enqueue_corr(b, g, s);
    if (!r) r = run_steps(L.post, s);
  }
  if (!r) r = run_steps(b->finish, s);
  return r;
}

}  // namespace

extern "C" {

int piv_batch_create(int device, int n_pairs, const piv_plan_desc* descs, piv_batch** out) {
  if (!out || !descs || n_pairs < 1) return fail(PIV_E_INVALID, "batch: bad argument");
  *out = nullptr;
  for (int i = 0; i < n_pairs; ++i) {
    if (descs[i].precision != PIV_FP32) return fail(PIV_E_INVALID, "batch: FP32 plans only");
    for (int k = 0; k < descs[i].n_passes && k < PIV_MAX_PASSES; ++k)
      if (descs[i].passes[k].fft_size > kMaxSmallFft)
        return fail(PIV_E_INVALID, "batch: FFT sizes up to %d (pair %d pass %d has %d)", kMaxSmallFft, i, k,
                    descs[i].passes[k].fft_size);
  }
  if (piv_device_count() <= 0) return fail(PIV_E_NO_DEVICE, "no HIP device visible");
  HIP_TRY(hipSetDevice(device));
  piv_batch* b = new piv_batch();
  b->device = device;
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
    batch_release(b);
    return fail(PIV_E_HIP, "batch: hipStreamCreate failed");
  }
  size_t max_passes = 0;
  for (int i = 0; i < n_pairs; ++i) {
    if (descs[i].in_dtype != descs[0].in_dtype) {
      batch_release(b);
      return fail(PIV_E_INVALID, "batch: every pair must share the input dtype");
    }
    piv_plan* P = nullptr;
    int r = piv_plan_create(device, &descs[i], &P);
    if (r) {
      batch_release(b);
      return r;
    }
    b->plans.push_back(P);
    max_passes = std::max(max_passes, P->passes.size());
  }
  b->levels.resize(max_passes);
  *out = b;
  return PIV_OK;
}

void piv_batch_destroy(piv_batch* b) { batch_release(b); }

int piv_batch_set_frames(piv_batch* b, const void* const* frames_a, const void* const* frames_b) {
  if (!b || !frames_a || !frames_b) return fail(PIV_E_INVALID, "batch: bad argument");
  HIP_TRY(hipSetDevice(b->device));
  HIP_TRY(hipStreamSynchronize(b->stream));
  for (size_t i = 0; i < b->plans.size(); ++i) {
    if (!frames_a[i] || !frames_b[i]) return fail(PIV_E_INVALID, "batch: null frame for pair %zu", i);
    b->plans[i]->ext_a = frames_a[i];
    b->plans[i]->ext_b = frames_b[i];
  }
  batch_drop_graph(b);
  int r = batch_build_tables(b);
  if (r) return r;
  b->frames_set = true;
  return PIV_OK;
}

int piv_batch_run(piv_batch* b) {
  if (!b || !b->frames_set) return fail(PIV_E_INVALID, "batch: frames not set");
  HIP_TRY(hipSetDevice(b->device));
  hipStream_t s = b->stream;
  // the frames may come from null-stream work (see piv_run)
  HIP_TRY(hipEventRecord(b->plans[0]->inputs_ready, nullptr));
  HIP_TRY(hipStreamWaitEvent(s, b->plans[0]->inputs_ready, 0));
  if (b->runs > 0) {
    if (!b->graph_exec) {
      HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      int r = batch_enqueue(b);
      hipGraph_t g = nullptr;
      hipError_t e = hipStreamEndCapture(s, &g);
      if (r) {
        if (g) (void)hipGraphDestroy(g);
        return r;
      }
      if (e != hipSuccess) return fail(PIV_E_HIP, "batch: graph capture failed: %s", hipGetErrorString(e));
      b->graph = g;
      HIP_TRY(hipGraphInstantiate(&b->graph_exec, b->graph, nullptr, nullptr, 0));
    }
    HIP_TRY(hipGraphLaunch(b->graph_exec, s));
  } else {
    int r = batch_enqueue(b);
    if (r) return r;
  }
  b->runs++;
  HIP_TRY(hipStreamSynchronize(s));
  return PIV_OK;
}

int piv_batch_tune(piv_batch* b, int repeats, int32_t* choices) {
  if (!b || !b->frames_set) return fail(PIV_E_INVALID, "batch: frames not set");
  HIP_TRY(hipSetDevice(b->device));
  if (repeats < 1) repeats = 1;
  hipStream_t s = b->stream;
  // one eager run fills every buffer the candidates read
  int r = batch_enqueue(b);
  if (r) return r;
  HIP_TRY(hipStreamSynchronize(s));
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  auto time_it = [&](auto&& launch) -> float {
    float best = 1e30f;
    for (int k = 0; k < repeats + 1; ++k) {
      (void)hipEventRecord(e0, s);
      if (launch()) return 1e30f;
      (void)hipEventRecord(e1, s);
      (void)hipEventSynchronize(e1);
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      if (k > 0) best = std::min(best, ms);  // the first repetition warms up
    }
    return best;
  };
  int ci = 0;
  for (size_t l = 0; l < b->levels.size(); ++l) {
    BatchLevel& L = b->levels[l];
    if (l > 0 && L.n_deform > 0) {
      int best_t = L.deform_threads;
      float best = 1e30f;
      for (int t : {128, 256, 512}) {
        L.deform_threads = t;
        const float ms = time_it([&] { return batch_enqueue_deform(b, L, s); });
        if (ms < best) {
          best = ms;
          best_t = t;
        }
      }
      L.deform_threads = best_t;
    }
    if (choices) choices[ci++] = L.deform_threads;
    for (auto& g : L.groups) {
      int best_x = g.xcd;
      float best = 1e30f;
      for (int x : {1, 0}) {
        g.xcd = x;
        const float ms = time_it([&] { return batch_enqueue_corr(b, g, s); });
        if (ms < best) {
          best = ms;
          best_x = x;
        }
      }
      g.xcd = best_x;
      if (g.line) {  // then the workgroup width of the line kernel
        int best_w = 0;
        best = 1e30f;
        for (int wi = 0; wi < 3; ++wi) {
          g.bw = kLineWidths[wi];
          g.d_prefix = g.d_prefix_bw[wi];
          g.total_blocks = g.total_blocks_bw[wi];
          const float ms = time_it([&] { return batch_enqueue_corr(b, g, s); });
          if (ms < best) {
            best = ms;
            best_w = wi;
          }
        }
        g.bw = kLineWidths[best_w];
        g.d_prefix = g.d_prefix_bw[best_w];
        g.total_blocks = g.total_blocks_bw[best_w];
      }
      if (choices) {
        choices[ci++] = g.xcd;
        choices[ci++] = g.bw;
      }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  batch_drop_graph(b);  // the next run captures the tuned shapes
  return PIV_OK;
}

int piv_batch_choice_count(const piv_batch* b, int32_t* count) {
  if (!b || !count) return fail(PIV_E_INVALID, "batch: bad argument");
  if (!b->frames_set) return fail(PIV_E_INVALID, "batch: frames not set");
  int c = 0;
  for (const auto& L : b->levels) c += 1 + 2 * (int)L.groups.size();
  *count = c;
  return PIV_OK;
}

int piv_batch_set_choices(piv_batch* b, const int32_t* choices, int count) {
  if (!b || !choices) return fail(PIV_E_INVALID, "batch: bad argument");
  if (!b->frames_set) return fail(PIV_E_INVALID, "batch: frames not set");
  int32_t want = 0;
  (void)piv_batch_choice_count(b, &want);
  if (count != want) return fail(PIV_E_INVALID, "batch: %d choices given, the batch takes %d", count, want);
  int ci = 0;
  for (const auto& L : b->levels) {  // check everything before changing anything
    const int t = choices[ci++];
    if (t != 128 && t != 256 && t != 512) return fail(PIV_E_INVALID, "batch: deformation block width %d", t);
    for (const auto& g : L.groups) {
      if (choices[ci] != 0 && choices[ci] != 1) return fail(PIV_E_INVALID, "batch: XCD flag %d", choices[ci]);
      const int w = choices[ci + 1];
      if (g.line ? (w != 1 && w != 2 && w != 4) : w != 1)
        return fail(PIV_E_INVALID, "batch: workgroup width %d", w);
      ci += 2;
    }
  }
  ci = 0;
  for (auto& L : b->levels) {
    L.deform_threads = choices[ci++];
    for (auto& g : L.groups) {
      g.xcd = choices[ci++];
      const int w = choices[ci++];
      if (g.line) {
        const int wi = line_width_index(w);
        g.bw = w;
        g.d_prefix = g.d_prefix_bw[wi];
        g.total_blocks = g.total_blocks_bw[wi];
      }
    }
  }
  batch_drop_graph(b);  // the next run captures the new shapes
  return PIV_OK;
}

int piv_device_arch(int device, char* name, int len) {
  if (!name || len < 1) return fail(PIV_E_INVALID, "bad argument");
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, device));
  snprintf(name, (size_t)len, "%s", prop.gcnArchName);
  return PIV_OK;
}

int piv_batch_pair_shape(const piv_batch* b, int pair, int32_t* m, int32_t* n) {
  if (!b || pair < 0 || pair >= (int)b->plans.size()) return fail(PIV_E_INVALID, "batch: pair out of range");
  const PassBuffers& L = b->plans[pair]->passes.back();
  if (m) *m = L.d.field_m;
  if (n) *n = L.d.field_n;
  return PIV_OK;
}

int piv_batch_get(const piv_batch* b, int pair, float* u, float* v, uint8_t* val_init, uint8_t* val_final,
                  int outputs_on_device) {
  if (!b || pair < 0 || pair >= (int)b->plans.size()) return fail(PIV_E_INVALID, "batch: pair out of range");
  HIP_TRY(hipSetDevice(b->device));
  const piv_plan* P = b->plans[pair];
  const PassBuffers& L = P->passes.back();
  const hipMemcpyKind k = outputs_on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
  if (u) HIP_TRY(hipMemcpyAsync(u, L.u, L.N * sizeof(float), k, b->stream));
  if (v) HIP_TRY(hipMemcpyAsync(v, L.v, L.N * sizeof(float), k, b->stream));
  if (val_init) HIP_TRY(hipMemcpyAsync(val_init, L.flags, L.N, k, b->stream));
  if (val_final) HIP_TRY(hipMemcpyAsync(val_final, P->val_final, L.N, k, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  return PIV_OK;
}

}  // extern "C"
