// Seeded synthetic particle-image pairs rendered on the GPU.
//
// The reference ships no generator; this is the device-side twin of
// openpiv_gpu/synthetic.py's particle model (same blob shape, cut-off,
// margin, symmetric placement, noise, rounding and clipping), with the
// host's PCG64 streams replaced by a counter-based hash so that every
// particle and pixel draws its own numbers in parallel:
//   u(key) = (splitmix64(seed * phi + key) >> 11) * 2^-53.
// Particle i uses keys 4i .. 4i+2 (x, y, I0) of stream 0; pixel p of frame f
// uses keys 2p, 2p+1 of stream 1 + f (Box-Muller).  Blobs are accumulated in
// 2^-12 fixed point with integer atomics, so the result does not depend on
// the order in which particles land.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "../../include/openpiv_hip.h"
#include "piv_kernels.h"

namespace piv {
int capi_fail(int code, const char* fmt, ...);
}

namespace {

constexpr double kFixScale = 4096.0;  // accumulator units per intensity count

__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

__device__ __forceinline__ double uniform01(uint64_t seed, uint64_t stream, uint64_t key) {
  const uint64_t h = splitmix64(seed * 0x9E3779B97F4A7C15ull ^ (stream << 58) ^ key);
  return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

__device__ void displacement(const piv_synth_field& f, int ht, double x, double y, double* du, double* dv) {
  if (f.kind == PIV_FIELD_UNIFORM) {
    *du = f.u0;
    *dv = f.v0;
    return;
  }
  if (f.kind == PIV_FIELD_SINE) {
    *du = f.amp_u * sin(2.0 * M_PI * y / f.period);
    *dv = f.amp_v * cos(2.0 * M_PI * x / f.period);
    return;
  }
  double u = f.gamma * (y - ht / 2.0), v = 0.0;
  for (int k = 0; k < f.n_vortices; ++k) {
    const double dx = x - f.cx[k], dy = y - f.cy[k];
    const double r2 = dx * dx + dy * dy;
    if (r2 > 0.0) {
      const double ut = f.circ[k] / (2.0 * M_PI) * (1.0 - exp(-r2 / (f.rc[k] * f.rc[k]))) / r2;
      u -= ut * dy;
      v += ut * dx;
    }
  }
  *du = u;
  *dv = v;
}

struct SynthArgs {
  piv_synth_field field;
  double sigma, i0_low, i0_high;
  int ht, wd, margin;
  long long n_particles;
  uint64_t seed;
  int32_t* acc_a;
  int32_t* acc_b;
  float* particles;  // optional n x 5: x, y, I0, du, dv
};

__device__ void splat(int32_t* acc, int ht, int wd, double xp, double yp, double i0, double sigma) {
  const int rad = (int)ceil(4.0 * sigma);
  const int cx = (int)rint(xp), cy = (int)rint(yp);
  const double lim = (4.0 * sigma) * (4.0 * sigma), inv2s2 = 1.0 / (2.0 * sigma * sigma);
  for (int oy = -rad; oy <= rad; ++oy) {
    const int py = cy + oy;
    if (py < 0 || py >= ht) continue;
    for (int ox = -rad; ox <= rad; ++ox) {
      const int px = cx + ox;
      if (px < 0 || px >= wd) continue;
      const double ddx = px - xp, ddy = py - yp, d2 = ddx * ddx + ddy * ddy;
      if (d2 > lim) continue;
      const double val = i0 * exp(-d2 * inv2s2);
      atomicAdd(&acc[(long)py * wd + px], (int32_t)llrint(val * kFixScale));
    }
  }
}

__global__ __launch_bounds__(256) void particle_kernel(SynthArgs a) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n_particles) return;
  const double m = a.margin;
  const double x = -m + (a.wd + 2.0 * m) * uniform01(a.seed, 0, 4 * (uint64_t)i);
  const double y = -m + (a.ht + 2.0 * m) * uniform01(a.seed, 0, 4 * (uint64_t)i + 1);
  const double i0 = a.i0_low + (a.i0_high - a.i0_low) * uniform01(a.seed, 0, 4 * (uint64_t)i + 2);
  double du, dv;
  displacement(a.field, a.ht, x, y, &du, &dv);
  splat(a.acc_a, a.ht, a.wd, x - du / 2.0, y - dv / 2.0, i0, a.sigma);
  splat(a.acc_b, a.ht, a.wd, x + du / 2.0, y + dv / 2.0, i0, a.sigma);
  if (a.particles) {
    float* p = a.particles + 5 * i;
    p[0] = (float)x;
    p[1] = (float)y;
    p[2] = (float)i0;
    p[3] = (float)du;
    p[4] = (float)dv;
  }
}

// Accumulator -> noisy, rounded, clipped output; clears the accumulator.
template <typename T>
__global__ __launch_bounds__(256) void finish_kernel(int32_t* acc, T* out, long npix, uint64_t seed, int frame,
                                                     float noise_sigma, float max_value) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npix) return;
  float val = (float)((double)acc[p] * (1.0 / kFixScale));
  acc[p] = 0;
  if (noise_sigma > 0.f) {
    const double u1 = uniform01(seed, 1 + frame, 2 * (uint64_t)p);
    const double u2 = uniform01(seed, 1 + frame, 2 * (uint64_t)p + 1);
    const double z = sqrt(-2.0 * log(1.0 - u1)) * cos(2.0 * M_PI * u2);
    val += noise_sigma * (float)z;
  }
  val = fminf(fmaxf(rintf(val), 0.f), max_value);
  out[p] = (T)val;
}

template <typename T>
hipError_t finish(int32_t* acc, void* out, long npix, uint64_t seed, int frame, float ns, float mx, hipStream_t s) {
  hipLaunchKernelGGL(finish_kernel<T>, dim3((unsigned)((npix + 255) / 256)), dim3(256), 0, s, acc, (T*)out, npix,
                     seed, frame, ns, mx);
  return hipGetLastError();
}

}  // namespace

extern "C" int piv_synth_pair(const piv_synth_field* field, const piv_synth_model* model, uint64_t seed, int height,
                              int width, void* frame_a, void* frame_b, int32_t* accum, float* particles,
                              void* stream) {
  using piv::capi_fail;
  if (!field || !model || !frame_a || !frame_b || !accum || height < 1 || width < 1)
    return capi_fail(PIV_E_INVALID, "piv_synth_pair: bad argument");
  if (field->kind < 0 || field->kind > 2 || field->n_vortices < 0 || field->n_vortices > PIV_SYNTH_MAX_VORTICES)
    return capi_fail(PIV_E_INVALID, "piv_synth_pair: bad field descriptor");
  if (!(model->sigma > 0) || model->sigma > 8 || model->margin < 0 || model->n_particles < 0 ||
      model->out_dtype < PIV_U8 || model->out_dtype > PIV_F32)
    return capi_fail(PIV_E_INVALID, "piv_synth_pair: bad particle model");
  const double peak = fmax(fabs(model->i0_low), fabs(model->i0_high));
  // a pixel sums at most the blobs within 4 sigma: keep the fixed-point sum in int32
  if (peak * kFixScale * (1.0 + 0.5 * 16.0 * model->sigma * model->sigma) > 2.0e9)
    return capi_fail(PIV_E_INVALID, "piv_synth_pair: intensities too large for the fixed-point accumulator");
  hipStream_t s = (hipStream_t)stream;
  SynthArgs a{};
  a.field = *field;
  a.sigma = model->sigma;
  a.i0_low = model->i0_low;
  a.i0_high = model->i0_high;
  a.ht = height;
  a.wd = width;
  a.margin = model->margin;
  a.n_particles = model->n_particles;
  a.seed = seed;
  const long npix = (long)height * width;
  a.acc_a = accum;
  a.acc_b = accum + npix;
  a.particles = particles;
  (void)hipGetLastError();
  if (a.n_particles > 0)
    hipLaunchKernelGGL(particle_kernel, dim3((unsigned)((a.n_particles + 255) / 256)), dim3(256), 0, s, a);
  hipError_t e = hipGetLastError();
  const float ns = (float)model->noise_sigma, mx = (float)model->max_value;
  for (int f = 0; f < 2 && e == hipSuccess; ++f) {
    int32_t* acc = f ? a.acc_b : a.acc_a;
    void* out = f ? frame_b : frame_a;
    switch (model->out_dtype) {
      case PIV_U8: e = finish<uint8_t>(acc, out, npix, seed, f, ns, mx, s); break;
      case PIV_U16: e = finish<uint16_t>(acc, out, npix, seed, f, ns, mx, s); break;
      default: e = finish<float>(acc, out, npix, seed, f, ns, mx, s); break;
    }
  }
  if (e != hipSuccess) return capi_fail(PIV_E_HIP, "piv_synth_pair: %s", hipGetErrorString(e));
  return PIV_OK;
}
