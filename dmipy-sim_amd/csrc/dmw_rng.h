// dmw_rng.h -- counter-based random numbers and the isotropic step direction.
//
// The reference draws, per sub-step, a JAX threefry key split and a normal(3)
// triple, and normalises it to a unit vector (dmipy_sim/physics.py:298-301); the
// only observable is an isotropic direction.  Here every walker owns a Philox
// stream (no per-walker key state in memory): the counter is
// (walker, instance, call, tag) and the key is the 64-bit seed.
//
// Philox4x32-10 follows Salmon, Moraes, Dror & Shaw, "Parallel random numbers:
// as easy as 1, 2, 3", SC'11: ten rounds of two 32x32->64 multiplies with the
// round multipliers 0xD2511F53 / 0xCD9E8D57 and the Weyl key increments
// 0x9E3779B9 / 0xBB67AE85.
//
// The direction is sampled with Archimedes' theorem (z uniform on [-1, 1],
// azimuth uniform) using only IEEE-exact operations (multiply, add, sqrt) and a
// Taylor polynomial for sin/cos(pi*f) on |f| <= 1/4, evaluated with explicit
// fused multiply-adds, so a CPU build and the GPU build produce identical bits.
#pragma once

#include <stdint.h>

#include "rn_arith.h"

#if defined(__HIPCC__)
#define DMW_HD __host__ __device__ __forceinline__
#else
#define DMW_HD static inline
#endif

namespace dmw {

struct u32x4 { uint32_t x, y, z, w; };

DMW_HD uint32_t mulhilo32(uint32_t a, uint32_t b, uint32_t* hi) {
  uint64_t p = (uint64_t)a * (uint64_t)b;
  *hi = (uint32_t)(p >> 32);
  return (uint32_t)p;
}

// a ^ b ^ c: one v_bitop3_b32 (truth table 0x96) on the GPU instead of two XORs.
DMW_HD uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
#else
  return a ^ b ^ c;
#endif
}

// Philox4x32-10 block: 4 x 32 random bits for counter c under key (k0, k1).
// Rounds 0 and 1 use plain XORs: with a counter whose words 1-3 and key are the
// same across a wavefront, round 0's first output word and round 1's first
// product stay in scalar registers (the bit-op builtin would move them to
// vector registers).
DMW_HD u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    uint32_t hi0, hi1;
    uint32_t lo0 = mulhilo32(0xD2511F53u, c.x, &hi0);
    uint32_t lo1 = mulhilo32(0xCD9E8D57u, c.z, &hi1);
    u32x4 n;
    n.x = r < 2 ? (hi1 ^ c.y ^ k0) : xor3(hi1, c.y, k0);
    n.y = lo1;
    n.z = r < 2 ? (hi0 ^ c.w ^ k1) : xor3(hi0, c.w, k1);
    n.w = lo0;
    c = n;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c;
}

// Uniform float in [0, 1) on a 2^-24 grid (exact conversion).
DMW_HD float u01(uint32_t b) { return (float)(b >> 8) * 0x1p-24f; }

// Counter tags separating the independent streams of one walker.
enum : uint32_t { TAG_STEP = 0u, TAG_INIT = 1u, TAG_POWLES = 2u };

// sin(pi f), cos(pi f) for |f| <= 1/4 from the Taylor series of sin and cos
// (coefficients (-1)^k pi^(2k+1)/(2k+1)! and (-1)^k pi^(2k)/(2k)!, rounded to
// float32); truncation error < 2e-9.
DMW_HD void sincospi_quarter(float f, float* s, float* c) {
  const float x = f * f;
  float ps = fmaf(x, 0x1.507834p-4f, -0x1.32d2ccp-1f);
  ps = fmaf(x, ps, 0x1.466bc6p+1f);
  ps = fmaf(x, ps, -0x1.4abbcep+2f);
  ps = fmaf(x, ps, 0x1.921fb6p+1f);
  *s = f * ps;
  float pc = fmaf(x, -0x1.a6d1f2p-6f, 0x1.e1f506p-3f);
  pc = fmaf(x, pc, -0x1.55d3c8p+0f);
  pc = fmaf(x, pc, 0x1.03c1f0p+2f);
  pc = fmaf(x, pc, -0x1.3bd3ccp+2f);
  *c = fmaf(x, pc, 1.0f);
}

// Unit vector uniformly distributed on the sphere from two 32-bit words.  The
// square root's argument is 0 or at least 2^-22 (z on a 2^-23 grid), inside the
// range where sqrt_rn equals sqrtf.
__device__ __forceinline__ void iso_direction(uint32_t a, uint32_t b, float* ux, float* uy, float* uz) {
  // z = 1 - 2 u01(a), t = 2 u01(b) and f = t - q/2 are exact, so their single
  // fused forms below round nothing (checked for every 24-bit input); fl(z z) <= 1,
  // so 1 - z z >= 0 needs no clamp.
  const float z = fmaf((float)(a >> 8), -0x1p-23f, 1.0f);   // in (-1, 1]
  const float s = sqrt_rn(1.0f - z * z);
  const float t = (float)(b >> 8) * 0x1p-23f;                // azimuth / pi in [0, 2)
  const float q = rintf(2.0f * t);                           // quadrant 0..4
  const float f = fmaf(q, -0.5f, t);                         // |f| <= 1/4
  float sf, cf;
  sincospi_quarter(f, &sf, &cf);
  // quadrant qi: (cos, sin)(pi t) = (cf, sf), (-sf, cf), (-cf, -sf), (sf, -cf) for
  // qi = 0..3 -- odd quadrants swap the pair, the sign of cos flips for qi = 1, 2
  // and that of sin for qi = 2, 3 (sign flips are exact; no branches)
  const uint32_t qi = (uint32_t)(int)q;
  const bool swap = (qi & 1u) != 0u;
  const float c0 = swap ? sf : cf, s0 = swap ? cf : sf;
  const float c = __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, c0) ^ (((qi + 1u) & 2u) << 30));
  const float sn = __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, s0) ^ ((qi & 2u) << 30));
  *ux = s * c;
  *uy = s * sn;
  *uz = z;
}

}  // namespace dmw
