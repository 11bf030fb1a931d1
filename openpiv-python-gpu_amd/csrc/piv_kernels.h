// Kernel launchers shared by the plan executor and the stage entry points.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace piv {

enum InDtype { kU8 = 0, kU16 = 1, kF32 = 2, kF64 = 3 };
enum SubpixelMethod { kGaussian = 0, kParabolic = 1, kCentroid = 2 };
// FFT sizes up to kMaxSmallFft keep the correlation plane on chip; larger
// ones (to kMaxFft) keep it in a device-memory slab per workgroup.
constexpr int kMaxSmallFft = 128;
constexpr int kMaxFft = 1024;

// Correlation of one pass: every window of the field is extracted from the
// two frames, mean-normalised, zero-padded to F x F, cross-correlated by FFT,
// and reduced to an fftshift-ed argmax plus a 3-point sub-pixel fit.
struct CorrArgs {
  const void* frame_a;
  const void* frame_b;
  int in_dtype;            // InDtype of both frames
  int ht, wd;              // frame shape
  int ss;                  // search window size
  int win_pad;             // (ss - interrogation window) / 2; frame A is zero outside
  int fsize;               // FFT size F (power of two, 16..128)
  int spacing, x_off, y_off;
  int m, n;                // field shape
  int normalize;           // subtract window means (forced on when padded)
  int subpixel;            // SubpixelMethod
  const float* up;         // predictor on this grid added to the result, or null
  const float* vp;
  float* u;
  float* v;
  int32_t* i_peak;         // optional integer peak row (fftshift-ed plane)
  int32_t* j_peak;         // optional integer peak column
  float* planes;           // optional N x F x F dump of the fftshift-ed planes
  int s2n_method;          // 0 peak2peak, 1 peak2mean, 2 peak2energy (gpu_process.py:1312-1365)
  int s2n_size;            // peak2peak exclusion half-width
  float* s2n;              // per-window signal-to-noise ratio, or null (not computed)
  void* scratch;           // F > kMaxSmallFft: correlate_big_scratch_bytes(F, 4, m * n) bytes
};

hipError_t launch_correlate(const CorrArgs& a, hipStream_t stream);
// One-lane-per-line variant for F = 16, 32 (search window <= F/2, no S2N, no
// shrunk interrogation window); launch_correlate picks it when it applies and
// is enabled (OPENPIV_HIP_CORR_LINE=0 in the environment turns it off).
bool correlate_line_applies(const CorrArgs& a);
bool correlate_line_enabled();
hipError_t launch_correlate_line(const CorrArgs& a, hipStream_t stream);

// FP64 correlation (dtype_f = float64): the same semantics with every value
// in double; frames of any InDtype (deformed frames are kF64).  scratch:
// correlate64_scratch_bytes(fsize) of device memory (planes of F = 128 live
// there instead of LDS).
struct CorrArgs64 {
  const void* frame_a;
  const void* frame_b;
  int in_dtype;
  int ht, wd;
  int ss, win_pad, fsize;
  int spacing, x_off, y_off;
  int m, n;
  int normalize;
  int subpixel;
  const double* up;
  const double* vp;
  double* u;
  double* v;
  int32_t* i_peak;
  int32_t* j_peak;
  double* planes;          // optional N x F x F fftshift-ed planes
  int s2n_method;
  int s2n_size;
  double* s2n;
  const double* twiddle;   // F/2 entries exp(-2 pi i k / F), (re, im) pairs
  double2* scratch;
};
size_t correlate64_scratch_bytes(int fsize);

// FFT sizes 256..1024 (both precisions): each workgroup keeps its F x F
// complex plane in a device-memory slab; correlate_big_blocks workgroups
// (capped by a 512 MiB slab budget) walk the windows.  FP32 takes its slabs
// from CorrArgs::scratch, FP64 from CorrArgs64::scratch; 0 bytes for F <= 128.
int correlate_big_blocks(int fsize, int real_bytes, int n_windows);
size_t correlate_big_scratch_bytes(int fsize, int real_bytes, int n_windows);
hipError_t launch_correlate_big(const CorrArgs& a, hipStream_t stream);
hipError_t launch_correlate_big64(const CorrArgs64& a, hipStream_t stream);
int correlate64_blocks(int fsize, int n_windows);
hipError_t launch_correlate64(const CorrArgs64& a, hipStream_t stream);

// One launch over the windows of many fields of the same FFT size and input
// dtype (device tables; every field's windows padded to whole blocks of
// correlate_windows_per_block(F)).  xcd: XCD-banded block order.
struct CorrBatch {
  const CorrArgs* fields;
  const int* block_prefix;  // n_fields entries: first block of each field
  int n_fields;
  int xcd;
};
int correlate_windows_per_block(int fsize);
// half: every field of the launch has search window <= F/2 (the single-pair
// launch of such a field uses the same arithmetic)
hipError_t launch_correlate_batch(const CorrBatch& b, int fsize, int in_dtype, int total_blocks, int max_ss,
                                  bool half, hipStream_t stream);
// bw: waves per workgroup of the batched line kernel (1, 2 or 4; a launch
// shape, the same arithmetic per window)
int correlate_line_windows_per_block(int fsize, int bw = 1);
hipError_t launch_correlate_line_batch(const CorrBatch& b, int fsize, int in_dtype, int total_blocks, int max_ss,
                                       int bw, hipStream_t stream);
size_t correlate_lds_bytes(int fsize, int ss);
// Residue-split variant for F = 64, 128 (search window <= F/2, no S2N, no
// shrunk interrogation window): every 1-D transform split by output index
// mod 4 over the four waves of a block.  launch_correlate picks it when it
// applies and is enabled (OPENPIV_HIP_CORR_SPLIT=0 turns it off).
bool correlate_split_applies(const CorrArgs& a);
bool correlate_split_enabled();
hipError_t launch_correlate_split(const CorrArgs& a, hipStream_t stream);
int correlate_split_windows_per_block(int fsize);
hipError_t launch_correlate_split_batch(const CorrBatch& b, int fsize, int in_dtype, int total_blocks,
                                        hipStream_t stream);
size_t correlate_split_lds_bytes(int fsize);

// Frame deformation: predictor bilinearly interpolated (with linear
// extrapolation) from the pass grid onto each pixel, then a backward bilinear
// warp of each frame at its own time step.  Output frames are FP32.
// R: the field / frame value type (float, or double for dtype_f = float64).
template <typename R>
struct DeformArgsT {
  const void* frame_a;
  const void* frame_b;
  int in_dtype;
  int ht, wd;
  const R* up;      // m x n predictor on the grid
  const R* vp;
  int m, n;
  R x0, y0;         // first grid node coordinates
  R spacing;        // grid spacing (pixels)
  R dt_a, dt_b;
  int warp_a, warp_b;  // 0 -> copy the frame unchanged
  R* out_a;
  R* out_b;
  const uint8_t* mask;  // frame mask (1 = masked: both outputs 0) or null
};
using DeformArgs = DeformArgsT<float>;
template <typename R>
hipError_t launch_deform(const DeformArgsT<R>& a, hipStream_t stream);
// Rows per thread of the FP32 deformation (a row group).
constexpr int kDeformRows = 4;
// One launch over the row groups of many frames (device tables); group r
// belongs to the last frame whose group_prefix entry is <= r.  threads: block
// width.
hipError_t launch_deform_batch(const DeformArgs* fields, const int* group_prefix, int n_fields, int total_groups,
                               int max_wd, int in_dtype, int threads, hipStream_t stream);

// Linear interpolation with extrapolation between two regular grids.
template <typename R>
struct GridInterpArgsT {
  const R* u_in;
  const R* v_in;
  int m_in, n_in;
  R x0_in, y0_in, sp_in;
  int m_out, n_out;
  R x0_out, y0_out, sp_out;
  R* u_out;
  R* v_out;
};
using GridInterpArgs = GridInterpArgsT<float>;
template <typename R>
hipError_t launch_grid_interp(const GridInterpArgsT<R>& a, hipStream_t stream);

// 3x3 NaN-aware neighbour-median test on u and v (flags OR-ed).
hipError_t launch_median_validate(const float* u, const float* v, int m, int n, float tol, uint8_t* flags,
                                  int32_t* count, hipStream_t stream);

// Largest neighbourhood half-width of the validation tests and of
// replacement: (2 * 8 + 1)^2 - 1 = 288 neighbours, whose statistics are found
// by rank counting (O(K^2) loads per vector for K neighbours); the spring
// system of half-width s splits into s^2 residue classes.
constexpr int kMaxNeighbourhoodSize = 8;

// Every validation test of ValidationGPU (gpu_validation.py:84-164), OR-ed:
// s2n < s2n_tol; |f - med| > median_tol; |f - med| > mad_tol * mad;
// |f - mean| > mean_tol over the (2 size + 1)^2 neighbourhood (centre
// excluded, gpu_validation.py:73-76, 194-205); and
// |f - mean(f)| > rms_tol * std(f) with the field statistics in rms
// (mean_u, std_u, mean_v, std_v, FP64).  A NaN tolerance turns a test off.
// only: validate where only[k] != 0 (others 0); extra: OR-ed in.
template <typename R>
struct ValidateArgsT {
  const R* u;
  const R* v;
  int m, n;
  int size;  // neighbourhood half-width, 1..kMaxNeighbourhoodSize
  const R* s2n;
  R s2n_tol, median_tol, mad_tol, mean_tol, rms_tol;
  const double* rms;
  const uint8_t* only;
  const uint8_t* extra;
  uint8_t* flags;
  int32_t* count;
  const uint8_t* fmask;  // field mask: never flagged (gpu_validation.py:207-210), or null
};
using ValidateArgs = ValidateArgsT<float>;
template <typename R>
hipError_t launch_validate(const ValidateArgsT<R>& a, hipStream_t stream);
// Field mean and population standard deviation of u and v (two-pass, FP64).
template <typename R>
hipError_t launch_field_stats(const R* u, const R* v, int count, double* stats, hipStream_t stream);

// Median (method 1) or mean (2) replacement of the flagged vectors from the
// (2 size + 1)^2 neighbourhood of (u_in, v_in) (ReplacementGPU,
// gpu_validation.py:272-290, 373-398); exclude: flagged neighbours left out
// (first iteration).  Writes u, v at flagged positions only; NaN where no
// neighbour is usable.
template <typename R>
hipError_t launch_neighbour_replace(const R* u_in, const R* v_in, R* u, R* v, const uint8_t* flags, int m, int n,
                                    int size, int method, int exclude, hipStream_t stream);
// Flagged vectors that came out NaN go back to (u0, v0) (ReplacementGPU.reset,
// gpu_validation.py:401-415); nan_flags[k] = 1 for them, else 0.
template <typename R>
hipError_t launch_reset_unresolved(R* u, R* v, const R* u0, const R* v0, const uint8_t* flags, uint8_t* nan_flags,
                                   int count, hipStream_t stream);

// Spring (Laplace) replacement of flagged vectors, solved on the device;
// each outlier is linked to the four vectors size rows / columns away
// (gpu_validation.py:279-283).
struct ReplaceWork {
  int32_t* slot;      // m*n
  int32_t* list;      // m*n
  double* work;       // 8 * m*n doubles
};
template <typename R>
hipError_t launch_spring_replace(R* u, R* v, const uint8_t* flags, const int32_t* count, int m, int n, int size,
                                 const ReplaceWork& w, hipStream_t stream);

// Unweighted DCT-PLS smoothing (GCV-chosen s) of u and v, in place.
template <typename R>
struct SmoothnWorkT {
  const R* cos_m;       // m x m orthonormal DCT-II matrix (row k = frequency)
  const R* cos_n;       // n x n
  const double* lambda; // m x n
  R* t0;                // 2 * m*n
  R* t1;                // 2 * m*n
  double* p_out;        // chosen log10(s)
  void* state;          // smoothn_state_bytes() of device scratch; its
                        // counter word must start (and stays) zero
};
using SmoothnWork = SmoothnWorkT<float>;
size_t smoothn_state_bytes();
template <typename R>
hipError_t launch_smoothn(R* u, R* v, int m, int n, double p_min, double p_max, double fixed_p, int auto_s,
                          const SmoothnWorkT<R>& w, hipStream_t stream);

// Weighted DCT-PLS smoothing of u and v in place for fields with masked or
// missing vectors (nearest-finite fill, weights, coarse start, relaxed
// iterations with a device-side convergence flag; masked vectors 0 on
// return).  m, n <= smoothn_masked_max_dim().
template <typename R>
struct SmoothnMaskedWorkT {
  const R* cos_m;       // m x m DCT-II matrix (row k = frequency)
  const R* cos_mt;      // its transpose
  const R* cos_n;
  const R* cos_nt;
  const double* lambda; // m x n
  const uint8_t* mask;  // m x n field mask (1 = masked) or null
  int32_t* nr;          // m*n
  R* f;                 // 2 m*n
  R* w;                 // m*n
  R* z;                 // 4 m*n (two slots)
  R* t0;                // 2 m*n
  R* t1;                // 2 m*n
  double* part;         // smoothn_masked_part_doubles(m)
  void* state;          // smoothn_masked_state_bytes()
  R* qbatch;            // 64 x 2 m*n when slow_gcv, else unused
  double* p_out;        // chosen log10(s) (may be null)
  int32_t* iters_out;   // iterations run (may be null)
  int slow_gcv;         // mean weight <= 0.9: GCV residual from the data misfit
};
using SmoothnMaskedWork = SmoothnMaskedWorkT<float>;
size_t smoothn_masked_state_bytes();
size_t smoothn_masked_part_doubles(int m);
int smoothn_masked_max_dim();
template <typename R>
hipError_t launch_smoothn_masked(R* u, R* v, int m, int n, double p_min, double p_max, double fixed_p, int auto_s,
                                 const SmoothnMaskedWorkT<R>& w, hipStream_t stream);

// Window correlation by DFT matrix products on MFMA (precision mode).
struct DftArgs {
  const float* win_a;   // n x ws x ws windows (FP32)
  const float* win_b;
  int n, ws, fsize;
  int precision;        // 0 fp16 split, 1 bf16 split, 2 fp16, 3 bf16
  int subpixel;
  const float* er;      // E = D_F[:, pad:pad+ws]: F x ws, real and imaginary parts
  const float* ei;
  const float* gr;      // G[x][k] = exp(+2 pi i k x / F): F x F
  const float* gi;
  float* workspace;     // dft_workspace_bytes(n, ws, fsize)
  float* u;
  float* v;
  int32_t* i_peak;
  int32_t* j_peak;
  float* planes;        // optional n x F x F fftshift-ed planes
};
size_t dft_workspace_bytes(int n, int ws, int fsize);
hipError_t launch_dft_correlate(const DftArgs& d, hipStream_t stream);

// Text describing the last failing launcher call on this thread ("" when the
// failure came from the kernel launch itself).
const char* launch_detail();
void set_launch_detail(const char* what, hipError_t e);

hipError_t launch_fill_i32(int32_t* p, int32_t value, int count, hipStream_t stream);
template <typename R>
hipError_t launch_scale(R* u, R* v, int count, R factor, hipStream_t stream);
// u = (u * mul) / div, v likewise (FP64 output scaling).
hipError_t launch_scale_div(double* u, double* v, int count, double mul, double div, hipStream_t stream);
template <typename R>
hipError_t launch_copy(R* dst, const R* src, int count, hipStream_t stream);
inline hipError_t launch_copy_f32(float* dst, const float* src, int count, hipStream_t stream) {
  return launch_copy<float>(dst, src, count, stream);
}
hipError_t launch_copy_u8(uint8_t* dst, const uint8_t* src, int count, hipStream_t stream);
// Several device-to-device copies (or zero fills: src == nullptr) in one launch.
constexpr int kCopyMax = 4;
struct CopySpec {
  void* dst;
  const void* src;
  size_t bytes;
};
struct CopyList {
  CopySpec c[kCopyMax];
};
hipError_t launch_copy_multi(const CopyList& l, int n, hipStream_t stream);
// Masked pixels of both frames set to 0 in place (mask: 16-byte aligned).
hipError_t launch_mask_frames(void* a, void* b, int in_dtype, const uint8_t* mask, size_t npx, hipStream_t stream);
// u = v = NaN and s2n = 0 (when non-null) at masked windows.
template <typename R>
hipError_t launch_apply_field_mask(R* u, R* v, R* s2n, const uint8_t* fmask, int count, hipStream_t stream);

}  // namespace piv
