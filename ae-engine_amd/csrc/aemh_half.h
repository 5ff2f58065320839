// aemh_half.h -- 16-bit storage formats of the mixed-precision mode.
// bfloat16: round-to-nearest-even on the top 16 bits of the FP32 pattern.
// float16: the hardware conversion (v_cvt_f16_f32, round to nearest even).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace aemh {

__device__ __forceinline__ uint16_t f32_to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);  // quiet NaN
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf16_to_f32(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ uint16_t f32_to_f16(float f) {
  _Float16 h = (_Float16)f;
  return __builtin_bit_cast(uint16_t, h);
}
__device__ __forceinline__ float f16_to_f32(uint16_t u) {
  return (float)__builtin_bit_cast(_Float16, u);
}

}  // namespace aemh
