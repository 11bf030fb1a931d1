/*
 * openpiv_hip.h -- C ABI of the MI355X (gfx950) PIV hot path.
 *
 * This is the drop-in boundary under the Python classes piv_gpu / PIVGPU
 * (openpiv_gpu/gpu_process.py of the reference).  The reference reaches the
 * GPU through CuPy: NVRTC kernels launched with RawModule.get_function()
 * (gpu_process.py:561-562, 1136-1145, 1184-1190) plus CuPy library calls
 * (cuFFT, reductions, nanmedian, spsolve, DCT).  Every entry point below
 * replaces one of those interfaces (cited per function) or the pass loop
 * that strings them together.
 *
 * Conventions
 *   - plain C types only; every function returns 0 on success and a negative
 *     PIV_E_* code on failure, with a message from piv_last_error()
 *     (thread-local);
 *   - "device" pointers are HIP device addresses on the plan's device; the
 *     caller owns all input and output buffers, the library owns plan
 *     workspaces;
 *   - a plan is not reentrant: one plan per host thread.
 */
#ifndef OPENPIV_HIP_H
#define OPENPIV_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PIV_MAX_PASSES 16

enum piv_status {
  PIV_OK = 0,
  PIV_E_INVALID = -1,     /* bad argument or unsupported setting */
  PIV_E_HIP = -2,         /* HIP runtime error (message has the HIP text) */
  PIV_E_NO_DEVICE = -3,   /* no usable GPU */
  PIV_E_NOMEM = -4
};

enum piv_in_dtype { PIV_U8 = 0, PIV_U16 = 1, PIV_F32 = 2, PIV_F64 = 3 /* FP64 plans only */ };
/* Processing precision, the reference's dtype_f (gpu_process.py:552-562):
 * FP64 computes every frame, window, plane and field value in double. */
enum piv_precision { PIV_FP32 = 0, PIV_FP64 = 1 };
enum piv_subpixel { PIV_GAUSSIAN = 0, PIV_PARABOLIC = 1, PIV_CENTROID = 2 };

/* One pass of the schedule.  Geometry is computed by the caller with the
 * closed forms of PIVFieldGPU (gpu_process.py:897-947); the library checks
 * it for consistency with the frame shape. */
typedef struct piv_pass_desc {
  int32_t search_size;        /* window size ss (gpu_process.py:675-680) */
  int32_t spacing;            /* ss - int(overlap_ratio * ss)           */
  int32_t x_offset, y_offset; /* centring offsets                        */
  int32_t field_m, field_n;   /* field shape                             */
  int32_t fft_size;           /* 2**ceil(log2(ss * n_fft)), 16..1024     */
  int32_t normalize;          /* subtract window means                   */
  int32_t subpixel;           /* enum piv_subpixel                       */
  float deforming_par;        /* gpu_process.py:1165-1171                */
  int32_t validate;           /* 1 -> median test with median_tol        */
  float median_tol;
  int32_t num_replacing_iters;/* 0 -> no replacement (at most 16)       */
  int32_t smooth;             /* smoothn after this pass                 */
  double smoothing_par;       /* <= 0 -> GCV-chosen                      */
  /* the other validation tests and replacement methods
   * (gpu_validation.py:84-164, 294-415; gpu_process.py:744-834); a NaN
   * tolerance (median_tol included) turns that test off */
  int32_t s2n_method;         /* enum piv_s2n (gpu_process.py:1312-1365)  */
  int32_t s2n_size;           /* peak2peak exclusion half-width           */
  float s2n_tol;              /* flag s2n < s2n_tol                       */
  float mad_tol;              /* flag |f - med| > mad_tol * MAD           */
  float mean_tol;             /* flag |f - mean| > mean_tol               */
  float rms_tol;              /* flag |f - mean(f)| > rms_tol * std(f)    */
  int32_t replacing_method;   /* enum piv_replacing                       */
  int32_t revalidate;         /* re-test replaced vectors each iteration  */
  int32_t window_size;        /* interrogation window inside the search
                               * window (shrink_ratio, gpu_process.py:654-662,
                               * 1066-1081); 0 = search_size            */
  int32_t validation_size;    /* half-width of the median / MAD / mean
                               * neighbourhood, (2 size + 1)^2 kernel
                               * (gpu_validation.py:73-76, 194-205);
                               * 0 = 1                                  */
  int32_t replacing_size;     /* half-width of the replacement kernel; the
                               * spring links reach size vectors away
                               * (gpu_validation.py:272-290); 0 = 1     */
  int32_t reserved;
} piv_pass_desc;

#define PIV_MAX_REPLACING_ITERS 16
enum piv_s2n { PIV_PEAK2PEAK = 0, PIV_PEAK2MEAN = 1, PIV_PEAK2ENERGY = 2 };
enum piv_replacing { PIV_SPRING = 0, PIV_MEDIAN = 1, PIV_MEAN = 2 };

typedef struct piv_plan_desc {
  int32_t height, width;      /* frame shape                             */
  int32_t in_dtype;           /* enum piv_in_dtype                       */
  int32_t n_passes;
  piv_pass_desc passes[PIV_MAX_PASSES];
  float out_scale;            /* scaling_par / dt (gpu_process.py:847-849) */
  int32_t use_graph;          /* capture the pass loop in a hipGraph     */
  int32_t precision;          /* enum piv_precision                      */
  int32_t reserved;
  double scaling_par;         /* FP64 plans: (u * scaling_par) / dt when */
  double dt;                  /* either is not 1 (out_scale is FP32's)   */
} piv_plan_desc;

typedef struct piv_plan piv_plan;

/* Error text of the last failing call on this thread. */
const char *piv_last_error(void);

/* Number of visible HIP devices (0 when none). */
int piv_device_count(void);

/* Builds a plan for one frame shape and pass schedule: allocates every
 * per-pass buffer and the smoothn tables.  Replaces PIVGPU.__init__ /
 * init_piv_fields and the two RawModule compilations
 * (gpu_process.py:465-562, 649-673). */
int piv_plan_create(int device, const piv_plan_desc *desc, piv_plan **out);
void piv_plan_destroy(piv_plan *plan);

/* Runs the whole pass loop on one image pair.  Replaces PIVGPU.__call__
 * (gpu_process.py:564-647) including the host-to-device upload (:582-583).
 * frame_a/frame_b: height*width elements of in_dtype, on the host
 * (frames_on_device == 0) or the device.  Outputs (final pass, field_m x
 * field_n of the last pass, row-major): u, v as float32 (float64 for FP64
 * plans); val_init = flags
 * before replacement (PIVGPU.init_val_locations, :770-771); val_final =
 * flags after replacement (PIVGPU.val_locations, :783-784).  Any output
 * pointer may be NULL.  outputs_on_device selects host or device outputs.
 * Device-resident frames are read after all work queued before the call on
 * the null stream.  Synchronous: returns after the outputs are written. */
int piv_run(piv_plan *plan, const void *frame_a, const void *frame_b, int frames_on_device, void *u, void *v,
            uint8_t *val_init, uint8_t *val_final, int outputs_on_device);

/* piv_run with device frames and device outputs that returns once the run is
 * queued on the plan's stream (no counterpart in the reference, whose
 * piv_gpu call is synchronous; for callers that stream pairs through one
 * plan).  Runs queued back to back execute in order; inputs must stay valid
 * and outputs and diagnostics are ready only after piv_plan_synchronize.
 * With a learned predictor the queued run is the learned path; the
 * fallback decision (gpu_process.py:619-628 has no counterpart; the plan's
 * max_reject_fraction) waits for piv_plan_synchronize, which reruns the full
 * schedule on the frames of every queued pair above the threshold, so its
 * frames must stay valid until then too. */
int piv_run_async(piv_plan *plan, const void *frame_a, const void *frame_b, void *u, void *v, uint8_t *val_init,
                  uint8_t *val_final);
/* Waits for every run queued on the plan's stream (and runs the fallbacks of
 * queued learned runs). */
int piv_plan_synchronize(piv_plan *plan);
/* Learned runs of the plan so far and how many of them fell back to the full
 * schedule (synchronous and queued runs; queued ones once synchronised). */
int piv_plan_learned_counts(const piv_plan *plan, int64_t *runs, int64_t *fallbacks);

/* Diagnostics of the last run: smoothing parameter log10(s) chosen after
 * pass k (NaN when no smoothing ran), and outlier counts before/after
 * replacement. */
int piv_plan_last_smoothing(const piv_plan *plan, int pass, double *log10_s);
int piv_plan_last_counts(const piv_plan *plan, int pass, int32_t *n_before, int32_t *n_after);

/* Frame mask (height x width bytes on the host, non-zero = masked) or NULL
 * to remove it.  Replaces the mask argument of PIVGPU (gpu_process.py:494,
 * 548-549) with its uses: frames zeroed under the mask (mask_frames,
 * :586, 682-687); per-pass field masks mask[int(y), int(x)] at the window
 * centres (PIVFieldGPU.get_field_mask, :919-920, 949-951); masked windows
 * give NaN displacement and S2N 0 (:1296-1310, 1338-1343); masked pixels of
 * deformed frames are 0 (cuda_interpolate_masked_frame, :1461-1506);
 * masked vectors are never flagged (gpu_validation.py:99-119, 207-210);
 * smoothing is the weighted one with masked vectors 0 on return
 * (SmoothnGPU with mask, :842; gpu_smoothn.py:123-297).  Not combinable
 * with a learned predictor. */
int piv_plan_set_mask(piv_plan *plan, const uint8_t *mask, int height, int width);
/* Iterations of the weighted smoothing after pass k in the last run (0 when
 * the pass was smoothed unweighted or not at all). */
int piv_plan_last_smoothing_iterations(const piv_plan *plan, int pass, int32_t *iterations);
/* Signal-to-noise ratio of the last pass of the last run (field_m x
 * field_n host values, float32 or float64 per the plan's precision) with
 * that pass's s2n_method / s2n_size, 0 at masked vectors.  Replaces
 * PIVGPU.s2n_ratio -> CorrelationGPU.s2n_ratio (gpu_process.py:864-867,
 * 1312-1376): recomputed on request from the frames the last pass read,
 * which the plan still holds.  Synchronises. */
int piv_plan_last_s2n(piv_plan *plan, void *s2n_out);

/* ---- stage entry points (device pointers, caller-owned; stream may be
 * NULL for the null stream) -------------------------------------------- */

/* Window extraction + normalisation + zero-padded FFT correlation + peak +
 * sub-pixel fit (+ the signal-to-noise ratio of pass->s2n_method into s2n
 * when s2n is non-NULL, CorrelationGPU.get_s2n_ratio, :1312-1365) for every
 * window of one pass.  Replaces cuda_get_stack
 * (gpu_process.py:1122-1147, 1420-1449) and CorrelationGPU.__call__'s
 * correlation chain (:1051-1097, 1201-1310).  up/vp (field_m x field_n) are
 * added to the result when non-NULL (:631-632).  i_peak/j_peak/planes are
 * optional (planes: N x F x F fftshift-ed correlation, float32). */
int piv_stage_correlate(const void *frame_a, const void *frame_b, int in_dtype, int height, int width,
                        const piv_pass_desc *pass, const float *up, const float *vp, float *u, float *v,
                        int32_t *i_peak, int32_t *j_peak, float *planes, float *s2n, void *stream);

/* FP64 variant of piv_stage_correlate (dtype_f = float64, USE_LONG,
 * gpu_process.py:561-562, 1378-1449): frames of any in_dtype (PIV_F64
 * included), predictor and outputs in double, planes N x F x F double.
 * Allocates its twiddles and scratch; synchronises the stream. */
int piv_stage_correlate64(const void *frame_a, const void *frame_b, int in_dtype, int height, int width,
                          const piv_pass_desc *pass, const double *up, const double *vp, double *u, double *v,
                          int32_t *i_peak, int32_t *j_peak, double *planes, double *s2n, void *stream);

/* Frame deformation: predictor (field_m x field_n on the pass grid)
 * interpolated onto every pixel with linear extrapolation, then the backward
 * bilinear warp of frame A at dt-1 and frame B at dt.  Replaces
 * CorrelationGPU.deform_frame + cuda_interpolate_frame
 * (gpu_process.py:1149-1192, 1508-1547).  Outputs are float32 frames. */
int piv_stage_deform(const void *frame_a, const void *frame_b, int in_dtype, int height, int width,
                     const piv_pass_desc *pass, const float *up, const float *vp, float *out_a, float *out_b,
                     void *stream);

/* Predictor interpolation from one pass grid to the next.  Replaces
 * PIVGPU.interpolate_displacement (gpu_process.py:730-742). */
int piv_stage_grid_interp(const piv_pass_desc *from, const piv_pass_desc *to, const float *u_in,
                          const float *v_in, float *u_out, float *v_out, void *stream);

/* 3x3 median validation.  Replaces ValidationGPU.median_validation
 * (gpu_validation.py:136-141, 166-210).  count (device int32) is
 * incremented by the number of flagged vectors. */
int piv_stage_median_validate(const float *u, const float *v, int m, int n, float median_tol, uint8_t *flags,
                              int32_t *count, void *stream);

/* Every validation test of the pass (median, MAD, mean, RMS; S2N when s2n
 * is non-NULL), OR-ed.  Replaces ValidationGPU.__call__
 * (gpu_validation.py:84-164).  only (optional): validate where only[k] != 0
 * (the complement of the reference's mask argument, :119, 222-226).
 * count (device int32) is incremented by the number of flags set. */
int piv_stage_validate(const float *u, const float *v, int m, int n, const float *s2n, const piv_pass_desc *pass,
                       const uint8_t *only, uint8_t *flags, int32_t *count, void *stream);

/* The replacement loop of one pass on u, v in place: pass->replacing_method,
 * pass->num_replacing_iters iterations, revalidation with the pass's tests
 * (s2n optional).  flags: in, the vectors to replace; out, the flags after
 * the loop (changed only when revalidating).  Replaces
 * PIVGPU.replace_outliers + ReplacementGPU (gpu_process.py:788-834;
 * gpu_validation.py:294-415).  Allocates its workspace; synchronises. */
int piv_stage_replace(float *u, float *v, uint8_t *flags, int m, int n, const piv_pass_desc *pass, const float *s2n,
                      void *stream);

/* piv_stage_replace that allocates nothing.  No counterpart in the
 * reference: for callers that cannot call hipMalloc (code under HIP graph
 * capture, tight serving loops, harnesses that forbid allocation).  Its
 * work buffers live in a fixed device array of the library (about 6 MB,
 * held by every GPU the library runs on until the process ends), so
 * m * n must not exceed PIV_NOALLOC_MAX_VECTORS (PIV_E_INVALID otherwise).
 * Queued on the stream, no synchronisation: the results are ready when the
 * stream's work is.  Every call on one GPU shares that GPU's array: calls
 * that overlap on one device (other streams, other threads) race. */
#define PIV_NOALLOC_MAX_VECTORS 65536
int piv_stage_replace_noalloc(float *u, float *v, uint8_t *flags, int m, int n, const piv_pass_desc *pass,
                              const float *s2n, void *stream);

/* Spring replacement of the flagged vectors in place.  Replaces
 * ReplacementGPU spring path + PIVGPU.replace_outliers
 * (gpu_validation.py:336-371, 381-415; gpu_process.py:788-834).  count is
 * the device int32 written by piv_stage_median_validate. */
int piv_stage_spring_replace(float *u, float *v, const uint8_t *flags, const int32_t *count, int m, int n,
                             void *stream);

/* smoothn of u and v in place; smoothing_par <= 0 selects the GCV search.
 * log10_s_out (device double, may be NULL) receives the chosen log10(s).
 * Replaces SmoothnGPU.__call__ (gpu_smoothn.py:123-231, 284-372). */
int piv_stage_smoothn(float *u, float *v, int m, int n, double smoothing_par, double *log10_s_out, void *stream);

/* Weighted smoothn of u and v in place: vectors that are non-finite or
 * under mask (device m x n bytes, may be NULL) get weight 0 and are filled
 * from their nearest finite neighbour; relaxed iterations to a relative
 * change of 1e-3 (at most 100), GCV re-run at iterations 2^j - 1; masked
 * vectors are 0 on return.  Falls back to piv_stage_smoothn's single pass
 * when every weight is 1.  iterations_out (device int32, may be NULL).
 * Replaces SmoothnGPU.__call__ with mask / missing data
 * (gpu_smoothn.py:123-297, 353-372).  m, n <= 1024.  Synchronises. */
int piv_stage_smoothn_masked(float *u, float *v, const uint8_t *mask, int m, int n, double smoothing_par,
                             double *log10_s_out, int32_t *iterations_out, void *stream);

/* Precision mode: correlation of window stacks by dense DFT matrix
 * products on the MFMA units (FP16/BF16 inputs, FP32 accumulation), beside
 * the FP32 FFT path of piv_stage_correlate.  win_a/win_b: n x ws x ws FP32
 * windows (device).  precision: 0 FP16 split hi/lo, 1 BF16 split hi/lo,
 * 2 FP16, 3 BF16.  Same normalisation, zero padding, fftshift, peak and
 * sub-pixel rules as gpu_process.py:1065-1310.  Allocates its workspace;
 * synchronizes the stream.  No counterpart in the reference (its only
 * correlation is the FP32/FP64 cuFFT path, gpu_process.py:1237-1246). */
int piv_stage_correlate_dft(const float *win_a, const float *win_b, int n_windows, int window_size, int fft_size,
                            int precision, int subpixel, float *u, float *v, int32_t *i_peak, int32_t *j_peak,
                            float *planes, void *stream);

/* ---- batches of uneven pairs -------------------------------------------
 * No counterpart in the reference, which processes one pair per
 * PIVGPU.__call__ (gpu_process.py:564-647).  A batch holds n_pairs plans
 * (each with its own frame shape and schedule, one shared input dtype).
 * Every pass level runs one deformation launch and one correlation launch
 * per FFT size over all pairs, then each pair's field stages; the loop is
 * one hipGraph after the first run.  Frames are caller-owned device
 * buffers registered once (read in place; re-register after moving them).
 * piv_batch_tune times the launch-shape candidates of every level on the
 * registered frames and keeps the fastest; choices (optional) receives, per
 * level, the deformation block width then, per correlation group, its
 * XCD-banded flag and its workgroup width in waves (1, 2 or 4 for the
 * one-lane-per-line kernel, 1 otherwise). */
typedef struct piv_batch piv_batch;
int piv_batch_create(int device, int n_pairs, const piv_plan_desc *descs, piv_batch **out);
void piv_batch_destroy(piv_batch *batch);
int piv_batch_set_frames(piv_batch *batch, const void *const *frames_a, const void *const *frames_b);
int piv_batch_run(piv_batch *batch);
int piv_batch_tune(piv_batch *batch, int repeats, int32_t *choices);
/* Number of tuning choices of the batch (one per level plus two per
 * correlation group), and applying a set of them -- e.g. from a tuning
 * cache keyed by the batch's description -- without timing anything
 * (PIV_E_INVALID on a wrong count or value, or before the frames are set;
 * nothing changes then). */
int piv_batch_choice_count(const piv_batch *batch, int32_t *count);
int piv_batch_set_choices(piv_batch *batch, const int32_t *choices, int count);
/* GPU architecture name of a device (e.g. "gfx950:sramecc+:xnack-"), for
 * keying tuning caches.  No counterpart in the reference. */
int piv_device_arch(int device, char *name, int len);
int piv_batch_pair_shape(const piv_batch *batch, int pair, int32_t *m, int32_t *n);
int piv_batch_get(const piv_batch *batch, int pair, float *u, float *v, uint8_t *val_init, uint8_t *val_final,
                  int outputs_on_device);

/* ---- learned coarse-displacement predictor -----------------------------
 * No counterpart in the reference: there the deformation predictor of a
 * pass is always the previous pass's field (gpu_process.py:620-625).  A
 * small encoder-decoder CNN maps a pair, box-averaged 8 x 8 and standardised
 * (piv_cnn_prepare_input), to a displacement field on nodes
 * (16 j + 8, 16 i + 8) of the frame; it can stand in for the coarse passes
 * of a plan (piv_plan_set_predictor).  Training and inference run on the
 * library's own kernels: BF16 weights and activations, FP32 accumulation,
 * FP32 gradients, Adam.  in_height/in_width: network input = frame / 8
 * (multiples of 8).  All calls on a network are ordered on the given stream;
 * the get_* calls synchronise the device. */
typedef struct piv_cnn piv_cnn;
int piv_cnn_create(int in_height, int in_width, int max_batch, piv_cnn **out);
void piv_cnn_destroy(piv_cnn *net);
/* desc: n_layers x 8 int32 (source buffer 0, source buffer 1 or -1, x2
 * upsampling of source 0, input channels from source 0 and 1, output
 * channels, stride, ReLU); offsets: n_layers x 2 (weight [co][3][3][ci] and
 * bias offsets in the flat FP32 parameter vector).  Buffer 0 is the input,
 * buffer l+1 the output of layer l. */
int piv_cnn_layout(const piv_cnn *net, int32_t *n_layers, int32_t *desc, int64_t *offsets, int64_t *n_params);
int piv_cnn_set_params(piv_cnn *net, const float *host_params);   /* also resets Adam */
int piv_cnn_get_params(const piv_cnn *net, float *host_params);
int piv_cnn_get_grads(const piv_cnn *net, float *host_grads);
/* activation (gradient = 0, BF16 widened; the last buffer is the FP32
 * planar output) or its gradient (NHWC FP32) of buffer 1..n_layers */
int piv_cnn_get_buffer(const piv_cnn *net, int buffer, int gradient, int batch, float *host_out);
/* input: device BF16 [batch][in_h][in_w][2]; out: device FP32
 * [batch][2][in_h/2][in_w/2] (u then v) or NULL */
int piv_cnn_forward(piv_cnn *net, const uint16_t *input, int batch, float *out, void *stream);
/* forward + weighted mean-square loss + backward; target: device FP32
 * [batch][2][in_h/2][in_w/2]; weight: device FP32 [batch][in_h/2][in_w/2]
 * or NULL; loss: device FP32 scalar or NULL.  Gradients stay in the net. */
int piv_cnn_gradients(piv_cnn *net, const uint16_t *input, const float *target, const float *weight, int batch,
                      float *loss, void *stream);
int piv_cnn_adam_step(piv_cnn *net, float lr, float beta1, float beta2, float eps, void *stream);
int piv_cnn_prepare_input(piv_cnn *net, const void *frame_a, const void *frame_b, int in_dtype, int height, int width,
                          uint16_t *out, void *stream);

/* Learned predictor for a plan: piv_run then evaluates the network on the
 * pair, interpolates its field to the grid of pass first_pass as that pass's
 * deformation predictor and runs passes first_pass..last only (first_pass
 * < 0: the last pass alone).  When the last pass's median test rejects more
 * than max_reject_fraction of the vectors, the same call runs the plan's
 * full schedule instead (fallback).  net = NULL removes the predictor.  The
 * frame must be 8 x (network input). */
int piv_plan_set_predictor(piv_plan *plan, piv_cnn *net, int first_pass, float max_reject_fraction);
int piv_plan_last_fallback(const piv_plan *plan, int32_t *fell_back, int32_t *rejected);

/* ---- synthetic particle-image pairs ------------------------------------
 * No counterpart in the reference (it ships no generator or tests).  The
 * device twin of openpiv_gpu/synthetic.py's particle model, used to build
 * training and evaluation streams without a host bottleneck: particles
 * uniform over the frame plus `margin` px, Gaussian blobs point-sampled at
 * pixel centres out to 4 sigma, frame A at p - d(p)/2 and frame B at
 * p + d(p)/2, additive Gaussian noise, rint, clip to [0, max_value].
 * Random numbers come from a counter-based hash of (seed, stream, index).
 * accum: 2 * height * width int32 device workspace, zero on entry (the call
 * leaves it zeroed).  particles: optional n_particles x 5 FP32 device
 * output (x, y, I0, du, dv). */
#define PIV_SYNTH_MAX_VORTICES 8
enum piv_field_kind { PIV_FIELD_UNIFORM = 0, PIV_FIELD_SINE = 1, PIV_FIELD_VORTICES = 2 };

typedef struct piv_synth_field {
  int32_t kind;               /* enum piv_field_kind                          */
  int32_t n_vortices;         /* <= PIV_SYNTH_MAX_VORTICES                    */
  double u0, v0;              /* uniform: (u, v)                              */
  double amp_u, amp_v, period;/* sine: u = amp_u sin(2 pi y / period), v = amp_v cos(2 pi x / period) */
  double gamma;               /* vortices: shear u += gamma (y - height / 2)  */
  double cx[PIV_SYNTH_MAX_VORTICES], cy[PIV_SYNTH_MAX_VORTICES];
  double rc[PIV_SYNTH_MAX_VORTICES], circ[PIV_SYNTH_MAX_VORTICES];  /* Lamb-Oseen core radius, circulation */
} piv_synth_field;

typedef struct piv_synth_model {
  double sigma, i0_low, i0_high, noise_sigma, max_value;
  int32_t out_dtype;          /* enum piv_in_dtype                            */
  int32_t margin;             /* px beyond each frame edge                    */
  int64_t n_particles;
} piv_synth_model;

int piv_synth_pair(const piv_synth_field *field, const piv_synth_model *model, uint64_t seed, int height, int width,
                   void *frame_a, void *frame_b, int32_t *accum, float *particles, void *stream);

/* ---- host input path for pair streams ----------------------------------
 * The reference's notebooks read frames with OpenCV imread(...,
 * IMREAD_ANYDEPTH) and upload them with cp.asarray (gpu_process.py:582-583).
 * piv_tiff_read decodes baseline uncompressed grayscale TIFF (8/16-bit
 * unsigned or 32-bit float samples, strips, either byte order) into caller
 * memory, e.g. a pinned buffer from piv_host_alloc; piv_memcpy_async moves
 * it on a copy stream; piv_plan_wait_event makes the plan's next piv_run
 * wait for an event recorded after that copy, so decoding, upload and
 * compute of consecutive pairs overlap. */
int piv_tiff_info(const char *path, int32_t *height, int32_t *width, int32_t *bits);
int piv_tiff_read(const char *path, void *out, size_t out_bytes);
int piv_host_alloc(void **ptr, size_t bytes);
int piv_host_free(void *ptr);
int piv_memcpy_async(void *dst, const void *src, size_t bytes, int kind, void *stream);
int piv_stream_wait_event(void *stream, void *event);
int piv_plan_wait_event(piv_plan *plan, void *event);

/* ---- device memory, streams and events --------------------------------
 * The reference allocates and moves data with CuPy (cp.asarray / cp.empty /
 * .get(), gpu_process.py:582-583, 1127, 316).  These entry points give a
 * caller without another GPU runtime the same services on the current
 * device (device 0 unless piv_set_device was called). */
int piv_set_device(int device);
int piv_malloc(void **ptr, size_t bytes);
int piv_free(void *ptr);
/* kind: 0 host->device, 1 device->host, 2 device->device */
int piv_memcpy(void *dst, const void *src, size_t bytes, int kind);
int piv_memset(void *ptr, int value, size_t bytes);
int piv_synchronize(void);
int piv_stream_create(void **stream);
int piv_stream_destroy(void *stream);
int piv_stream_synchronize(void *stream);
int piv_event_create(void **event);
int piv_event_destroy(void *event);
int piv_event_record(void *event, void *stream);
int piv_event_elapsed_ms(void *start, void *stop, float *ms);

#ifdef __cplusplus
}
#endif

#endif /* OPENPIV_HIP_H */
