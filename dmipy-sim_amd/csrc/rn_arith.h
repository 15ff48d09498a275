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
//                exponent stays inside [-100, +100]; x == +-0 keeps the IEEE sign.
//
// For those inputs they execute the compiler's own instruction sequence minus the
// steps that are identities there, and they are checked against sqrtf (every float in
// the range) and against IEEE division (random operand pairs) on the GPU by
// tests/gpu_selftest/arith_check.hip.
#pragma once

#include <hip/hip_runtime.h>

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

}  // namespace dmw
