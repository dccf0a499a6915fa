// Correctly rounded quotients a_k / d that share one reciprocal of d.
//
// IEEE division costs about ten VALU instructions and one reciprocal per
// quotient.  When several numerators share one divisor, this computes instead
//   y  = RN(1/d)              hardware estimate y0 + one fused Newton step
//   q  = RN(a*y)
//   r  = a - d*q              exact (one fma)
//   q' = RN(q + r*y)          = RN(a/d)              (Markstein's correction)
// three instructions per quotient after the reciprocal.
//
// Where it is exact (div_rn_ok):
// * d in [2^-78, 2^40).  For every such d and every estimate y0 within one
//   ulp of 1/d (the documented accuracy of v_rcp_f32), one Newton step gives
//   RN(1/d), except for an all-ones significand when y0 is the lower of the
//   two candidates (the iterate lands on a midpoint and ties to even); the
//   hardware estimate is the upper one for every all-ones d in range, so
//   rcp_rn(d) = RN(1/d) throughout.
// * every numerator zero or of magnitude at least 2^-80.  Then q, r and q'
//   stay in the normal range (|q| >= 2^-120) and r is exact, so each step is
//   invariant under scaling a and d by powers of two, and the correction is
//   exact when it is exact for every pair of significands a, d in [1, 2).
//   A zero numerator gives +0 (IEEE gives the numerator's sign).
// tests/div_rn_check.c checks the reciprocal step on the CPU for every d in
// range and every estimate; tests/div_rn_check.hip checks the reciprocal on
// the device for every d in range and the quotient for every pair of
// significands.
//
// Numerators that are differences of coordinates need no check when every
// coordinate is 0 or of magnitude at least 2^-56: both are then multiples of
// 2^-79, so a nonzero difference is at least 2^-79 (div_rn_ok_d).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace lps_div {

__device__ __forceinline__ float rcp_rn(float d) {
    const float y0 = __builtin_amdgcn_rcpf(d);
    const float e = __builtin_fmaf(-d, y0, 1.f);
    return __builtin_fmaf(e, y0, y0);
}

__device__ __forceinline__ float div_rn(float a, float d, float y) {
    const float q = a * y;
    const float r = __builtin_fmaf(-q, d, a);
    return __builtin_fmaf(r, y, q);
}

// 2 * bits(2^-80) - 1: the smallest value of 2 * bits(a) - 1 (sign dropped,
// zero wrapping to the top) for |a| >= 2^-80.
constexpr uint32_t kNumMin = 2u * 0x17800000u - 1u;

__device__ __forceinline__ uint32_t num_key(float a) { return (__float_as_uint(a) << 1) - 1u; }

// d >= 2^-78 and every numerator 0 or >= 2^-80 (keys from num_key()); the
// upper bound d < 2^40 is the caller's (a uniform bound on d).
__device__ __forceinline__ bool div_rn_ok(float d, uint32_t key_min) { return d >= 0x1p-78f && key_min >= kNumMin; }

// The same when the numerators are known to be 0 or >= 2^-80.
__device__ __forceinline__ bool div_rn_ok_d(float d) { return d >= 0x1p-78f; }

// Whether a point has a coordinate that breaks the no-check rule for
// differences (0 < |c| < 2^-56), with the keys of num_key().
__device__ __forceinline__ bool tiny_coord(float x, float y, float z) {
    const uint32_t k = min(min(num_key(x), num_key(y)), num_key(z));
    return k < 2u * 0x23800000u - 1u;  // 2 * bits(2^-56) - 1
}

}  // namespace lps_div
