// The warp sampler's float arithmetic in forms cheaper than the plain operators and equal to them bit for bit on the
// inputs the sampler gives them; tests/devcheck/sample_math.hip runs each against the plain form on the device.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace lz {

// Grey level g / 255, correctly rounded: the product by the rounded reciprocal, corrected by one fused multiply-add
// (the residual g - 255 q is exact in it). Equal to the IEEE quotient for all 256 levels.
__device__ __forceinline__ float grey(uint32_t g) {
  constexpr float c = 1.0f / 255.0f;
  const float f = static_cast<float>(g), q = __fmul_rn(f, c);
  return __fmaf_rn(__fmaf_rn(-q, 255.0f, f), c, q);
}

// 1 / d correctly rounded: v_rcp_f32 and one Newton step of fused multiply-adds equal the IEEE quotient for every
// normal d with |d| < 2^126, both signs; the rest (zero, subnormal, |d| >= 2^126, inf, NaN) take the IEEE division.
__device__ __forceinline__ float recip_rn(float d) {
  const float a = fabsf(d);
  if (a >= 0x1p-126f && a < 0x1p126f) {
    const float r = __builtin_amdgcn_rcpf(d);
    return __fmaf_rn(__fmaf_rn(-d, r, 1.0f), r, r);
  }
  return __fdiv_rn(1.0f, d);
}

// x - (float)(int) x for x >= 0 below 2^31, as one v_fract_f32 (exact there: the difference is representable).
__device__ __forceinline__ float frac_pos(float x) { return __builtin_amdgcn_fractf(x); }

}  // namespace lz
