// Buffer-resource loads and stores with 32-bit offsets (gfx950 raw buffer
// instructions): the base address lives in scalar registers, each lane
// supplies a 32-bit byte offset and a scalar offset can select a row, so an
// access needs no 64-bit address arithmetic.  Offsets past the record count
// read 0 (the callers stay in range).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace piv {

using rsrc_t = __amdgpu_buffer_rsrc_t;

// Raw (unstructured) buffer over bytes [p, p + bytes); bytes < 2^31.
__device__ __forceinline__ rsrc_t make_rsrc(const void* p, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, 0x00020000);
}

__device__ __forceinline__ float buf_f32(rsrc_t r, uint32_t voff, uint32_t soff) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0));
}

// Pixel idx of a frame of T (uint8_t, uint16_t or float) as float.
template <typename T>
__device__ __forceinline__ float buf_px(rsrc_t r, uint32_t idx) {
  if constexpr (sizeof(T) == 4)
    return buf_f32(r, idx * 4u, 0u);
  else if constexpr (sizeof(T) == 2)
    return (float)(uint32_t)__builtin_amdgcn_raw_buffer_load_b16(r, (int)(idx * 2u), 0, 0);
  else
    return (float)(uint32_t)__builtin_amdgcn_raw_buffer_load_b8(r, (int)idx, 0, 0);
}

// Whether a frame of npx pixels of size es fits one raw buffer.
__host__ __device__ __forceinline__ bool buffer_fits(uint64_t npx, uint64_t es) { return npx * es < (1ull << 31); }

}  // namespace piv
