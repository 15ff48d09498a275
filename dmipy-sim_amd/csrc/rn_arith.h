// rn_arith.h -- correctly rounded float32 square root and quotient for the walk's
// operand ranges, without the range scaling of the general IEEE sequences.
//
// The compiler expands sqrtf and '/' (correctly rounded, as the reference's float32
// arithmetic is) into the hardware approximation followed by exact refinement, plus
// steps that only matter near the ends of the float range: input scaling by 2^32 for
// x < 2^-96 and a zero/infinity class fix-up (sqrt); v_div_scale on both operands,
// v_div_fmas and v_div_fixup (division).  The walk's operands never reach those ends
// -- squared lengths of steps and positions, and quotients of comparable lengths or of
// a length by its norm -- so these forms keep the refinement and drop the rest:
//
//   sqrt_rn(x)   exact (equal to sqrtf) for x == +0 and for finite x >= 2^-96;
//   div_rn(x, y) exact (equal to x / y) for finite x, normal y, and a quotient whose
//                exponent stays inside [-100, +100]; x == +-0 keeps the IEEE sign;
//   div_rn_r(x, y, r)  the same quotient for a divisor known in advance, given its
//                correctly rounded reciprocal r = RN(1/y): the hardware reciprocal
//                and its Newton step are skipped.  q0 = x r is within two ulps of
//                x/y, one residual step makes it faithful, and the last step rounds
//                correctly (Markstein: faithful q, exact residual x - y q, r = RN(1/y)).
//
// For those inputs sqrt_rn and div_rn execute the compiler's own instruction sequence
// minus the steps that are identities there.  All three are checked against sqrtf
// (every float in the range) and against IEEE division (random operand pairs) on the
// GPU by tests/gpu_selftest/arith_check.hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#pragma clang fp contract(off)

namespace dmw {

__device__ __forceinline__ float sqrt_rn(float x) {
  const float s = __builtin_amdgcn_sqrtf(x);                       // within 1 ulp
  const int si = __builtin_bit_cast(int, s);
  const float down = __builtin_bit_cast(float, si - 1);
  const float up = __builtin_bit_cast(float, si + 1);
  const float r_down = __builtin_fmaf(-down, s, x);                // x - down * s
  const float r_up = __builtin_fmaf(-up, s, x);                    // x - up * s
  float r = r_down <= 0.0f ? down : s;
  r = r_up > 0.0f ? up : r;
  return r;
}

__device__ __forceinline__ float div_rn(float x, float y) {
  const float a = __builtin_amdgcn_rcpf(y);                        // within 1 ulp
  const float e = __builtin_fmaf(-y, a, 1.0f);
  const float r = __builtin_fmaf(e, a, a);
  const float q0 = x * r;
  const float m0 = __builtin_fmaf(-y, q0, x);
  const float q1 = __builtin_fmaf(m0, r, q0);
  const float m1 = __builtin_fmaf(-y, q1, x);
  const float q = __builtin_fmaf(m1, r, q1);
  return x == 0.0f ? q0 : q;                                       // +-0 with the IEEE sign
}

__device__ __forceinline__ float div_rn_r(float x, float y, float r) {
  const float q0 = x * r;
  const float m0 = __builtin_fmaf(-y, q0, x);
  const float q1 = __builtin_fmaf(m0, r, q0);
  const float m1 = __builtin_fmaf(-y, q1, x);
  const float q = __builtin_fmaf(m1, r, q1);
  return x == 0.0f ? q0 : q;                                       // +-0 with the IEEE sign
}

// Host: RN(1/y) for div_rn_r.  1/y in float64 rounded to float32 can miss by one
// ulp when 1/y lies within 2^-53 of a float32 midpoint, so the neighbour with the
// smallest residual |1 - y r| is taken (y r is exact in float64: 24 + 24 bits).
inline float recip_rn(float y) {
  const float r0 = (float)(1.0 / (double)y);
  float best = r0;
  double res = std::fabs(1.0 - (double)y * (double)r0);
  for (const float c : {std::nextafter(r0, -INFINITY), std::nextafter(r0, INFINITY)}) {
    const double rc = std::fabs(1.0 - (double)y * (double)c);
    if (rc < res) {
      best = c;
      res = rc;
    }
  }
  return best;
}

// Host: the least float s >= 0 with sqrtf(s) >= r (r > 0 finite), so that for every
// float s >= 0, sqrt_rn(s) >= r exactly when s >= sqrt_cut(r) (square root correctly
// rounded and monotone; host sqrtf is the IEEE square root).
inline float sqrt_cut(float r) {
  float s = r * r;
  while (s > 0.0f && std::sqrt(std::nextafter(s, 0.0f)) >= r) s = std::nextafter(s, 0.0f);
  while (std::sqrt(s) < r) s = std::nextafter(s, INFINITY);
  return s;
}

}  // namespace dmw
