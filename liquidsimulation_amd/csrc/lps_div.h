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
// * d in [2^-78, 2^40) and d's significand not all ones.  For every such d
//   and every estimate y0 within one ulp of 1/d (the documented accuracy of
//   v_rcp_f32), one Newton step gives RN(1/d); the all-ones significand is
//   the exception (the iterate lands on a midpoint and ties to even), so it is
//   left to IEEE division.
// * every numerator zero or of magnitude at least 2^-80.  Then q, r and q'
//   stay in the normal range (|q| >= 2^-120) and r is exact, so each step is
//   invariant under scaling a and d by powers of two, and the correction is
//   exact when it is exact for every pair of significands a, d in [1, 2).
//   A zero numerator gives +0 (IEEE gives the numerator's sign).
// tests/div_rn_check.c checks the reciprocal step on the CPU for every d in
// range and every estimate; tests/div_rn_check.hip checks the reciprocal on
// the device for every d in range and the quotient for every pair of
// significands.
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

// d in [2^-78, 2^40) with a significand other than all ones; the upper bound
// is the caller's (a uniform bound on d), the numerators' keys are num_key().
__device__ __forceinline__ bool div_rn_ok(float d, uint32_t key_min) {
    return d >= 0x1p-78f && ((__float_as_uint(d) + 1u) & 0x7fffffu) != 0u && key_min >= kNumMin;
}

}  // namespace lps_div
