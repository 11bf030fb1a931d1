// Wave-level reductions for gfx950 (wave64) shared by the field kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace piv {

// A double moved across lanes by one DPP control (both 32-bit halves);
// lanes of rows outside ROW_MASK read 0.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_d(double v) {
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, CTRL, ROW_MASK, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), CTRL, ROW_MASK, 0xF, false);
  return __builtin_bit_cast(double, ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

// Sum over the 64 lanes of a full wave, returned to every lane: DPP steps
// within quads (lane ^ 1, ^ 2), half rows and rows (mirrors), then the row
// broadcasts of lanes 15 and 31 carry the partial sums up to lane 63, which
// is read back as a wave-uniform value.  Fixed order: deterministic.
__device__ __forceinline__ double wave_sum(double x) {
  x += dpp_d<0xB1, 0xF>(x);   // quad_perm [1,0,3,2]
  x += dpp_d<0x4E, 0xF>(x);   // quad_perm [2,3,0,1]
  x += dpp_d<0x141, 0xF>(x);  // row_half_mirror
  x += dpp_d<0x140, 0xF>(x);  // row_mirror
  x += dpp_d<0x142, 0xA>(x);  // row_bcast15 into rows 1 and 3
  x += dpp_d<0x143, 0xC>(x);  // row_bcast31 into rows 2 and 3
  const uint64_t u = __builtin_bit_cast(uint64_t, x);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, 63);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), 63);
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

}  // namespace piv
