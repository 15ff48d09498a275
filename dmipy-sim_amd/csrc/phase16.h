// phase16.h -- host-side tables of the matrix-core phase update.
//
// The walk kernel's phase product runs on v_mfma_f32_16x16x16f16 with both
// operands split into float16 hi + lo parts.  The waveform side is prepared
// here once per plan: G * g_scale as hi/lo float16 quadruples (x, y, z, 0), the
// column sums of G between emission points (the phase of a walker resting at the
// start point), and the
// power-of-two scales that keep either operand at most 2^14 in magnitude (a
// factor of four below the float16 limit, 65504).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace dmw {

constexpr double kHalfOperandMax = 16384.0;

// Largest power of two not above x, clamped to [2^-60, 2^60]; 1 for x <= 0, inf or NaN.
inline double pow2_floor(double x) {
  if (!(x > 0.0) || !std::isfinite(x)) return 1.0;
  int e = 0;
  (void)std::frexp(x, &e);  // x = f 2^e, f in [0.5, 1)
  return std::ldexp(1.0, std::max(-60, std::min(60, e - 1)));
}

inline uint16_t half_bits(double v) {
  const _Float16 h = (_Float16)v;  // round to nearest even
  uint16_t b;
  std::memcpy(&b, &h, 2);
  return b;
}

// Scale of the waveform operand: max |g| * scale <= 2^14.
inline double phase_g_scale(const float* g, size_t n) {
  double m = 0.0;
  for (size_t i = 0; i < n; ++i) m = std::max(m, std::fabs((double)g[i]));
  return m > 0.0 ? pow2_floor(kHalfOperandMax / m) : 1.0;
}

// Scale of the walker operand.  Per sub-step a walker moves at most its step
// (reflections, mirrors and folds do not lengthen it; the exterior mirror of the
// packed myelin walk at most triples it; clamps move it by a fraction of a step),
// so n_sub times the displacement from the start stays below
// 4 n_sub (n_t n_sub) step_max.
inline double phase_p_scale(int n_t, int n_sub, double step_max, double factor) {
  const double bound = 4.0 * (double)n_sub * (double)n_t * (double)n_sub * step_max * std::fabs(factor);
  return bound > 0.0 ? pow2_floor(kHalfOperandMax / bound) : 1.0;
}

// g (n_t, Mp, 3) -> out: float16 hi and lo parts of g * scale, four per entry
// (x, y, z, 0), laid out as the kernel's 16-byte per-lane chunks:
// chunk ((t * 2 + part) * Mp32 / 32 + m / 32) * 16 + m % 16 holds measurements m
// (column tile 2p, first 8 bytes) and m + 16 (tile 2p + 1); Mp32 = Mp rounded up to 32.
inline void phase_half_table(const float* g, int n_t, int Mp, double scale, std::vector<uint2>& out) {
  const int Mp32 = (Mp + 31) / 32 * 32;
  out.assign((size_t)n_t * 2 * Mp32, make_uint2(0u, 0u));
  for (int t = 0; t < n_t; ++t)
    for (int m = 0; m < Mp; ++m) {
      uint16_t hi[4] = {0, 0, 0, 0}, lo[4] = {0, 0, 0, 0};
      for (int c = 0; c < 3; ++c) {
        const double v = (double)g[((size_t)t * Mp + m) * 3 + c] * scale;
        hi[c] = half_bits(v);
        _Float16 h;
        std::memcpy(&h, &hi[c], 2);
        lo[c] = half_bits(v - (double)h);
      }
      for (int part = 0; part < 2; ++part) {
        const uint16_t* q = part ? lo : hi;
        const size_t chunk = ((size_t)t * 2 + part) * (Mp32 / 32) + m / 32;
        out[(chunk * 16 + m % 16) * 2 + (m / 16) % 2] = make_uint2(q[0] | ((uint32_t)q[1] << 16), q[2]);
      }
    }
}

// Column sums of g over the segments between emission points: segment 0 covers
// samples [0, e_0], segment k (e_{k-1}, e_k], the last (e_{n_echo-1}, n_t - 1]
// (one segment, the whole walk, without echoes).  out (n_echo + 1, 3, Mp),
// float64 accumulation rounded once.
inline void phase_column_sums(const float* g, int n_t, int Mp, const int* echo_idx, int n_echo,
                              std::vector<float>& out) {
  out.assign((size_t)(n_echo + 1) * 3 * Mp, 0.0f);
  int t0 = 0;
  for (int k = 0; k <= n_echo; ++k) {
    const int t1 = k < n_echo ? echo_idx[k] : n_t - 1;
    for (int m = 0; m < Mp; ++m)
      for (int c = 0; c < 3; ++c) {
        double s = 0.0;
        for (int t = t0; t <= t1; ++t) s += (double)g[((size_t)t * Mp + m) * 3 + c];
        out[((size_t)k * 3 + c) * Mp + m] = (float)s;
      }
    t0 = t1 + 1;
  }
}

}  // namespace dmw
