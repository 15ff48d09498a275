// screen.h -- conservative wall screen of the packed-substrate walk.
//
// The reference tests every object of a packing on every sub-step
// (geometries.py:2067-2093 for PackedCylinders.permeate, 2667-2691 for
// PackedSpheres): ray against circle/sphere, first valid root in (EPS, |s|).
// Almost every sub-step is far from every wall, and then its result is r + s
// exactly (geometries.py:2099-2102).  The screen decides, from squared distances
// only, that a sub-step cannot produce a valid root for a cached object; only
// the walkers it cannot clear run the reference's arithmetic.
//
// For an object of radius R centred at c, f(x) = |x - c|^2 - R^2 along the
// segment x(tau) = p + tau s, tau in [0, 1], is a convex parabola with leading
// coefficient |s|^2, so
//   max_tau f <= max(f(p), f(p + s))                      (convexity), and
//   min_tau f >= min(f(p), f(p + s)) - |s|^2 / 4           (vertex depth).
// The segment therefore stays at least a margin m outside the wall when
// min(f0, f1) > 2 R m + m^2 + |s|^2 / 4, and inside it when max(f0, f1) <
// m^2 - 2 R m.  In either case the reference's root test finds no valid root
// for that object: its float32 roots and the screen's float32 values differ
// from exact geometry by far less than m, which the host sizes from the cell
// side, the largest radius and the step (screen_margin below; the bound is
// exercised against the oracle's float32 root test on adversarial configurations
// in tests/test_screen.py).
#pragma once

#include <math.h>

#if defined(__HIPCC__)
#define DMW_SCR_HD __host__ __device__ __forceinline__
#else
#define DMW_SCR_HD static inline
#endif

namespace dmw {

// f = |x - c|^2 - R^2 in float32 (dim 2 ignores z).
DMW_SCR_HD float screen_f(float px, float py, float pz, float cx, float cy, float cz, float R, int dim) {
  const float ex = px - cx, ey = py - cy;
  float d2 = fmaf(ex, ex, ey * ey);
  if (dim == 3) {
    const float ez = pz - cz;
    d2 = fmaf(ez, ez, d2);
  }
  return fmaf(-R, R, d2);
}

// True when the segment from the point with value f0 to the point with value f1
// keeps at least the margin from the wall of radius R:
//   m2 = 2 m, c_out = m^2 + |s|max^2 / 4, c_in = m^2.
DMW_SCR_HD bool screen_clear(float f0, float f1, float R, float m2, float c_out, float c_in) {
  const float t_out = fmaf(R, m2, c_out);
  const float t_in = fmaf(-R, m2, c_in);
  return (fminf(f0, f1) > t_out) | (fmaxf(f0, f1) < t_in);
}

// Cell lists of the screened walk: a cell's list start carries, in the bits from
// SCR_LEVEL_SHIFT up, the level l of the box its list was built for: the cell
// expanded by (1 - l/2) times the walk's hysteresis width on every side (l = 0,
// 1, 2; the grid builder shrinks the box of a cell whose list would not fit the
// walker's object cache).
constexpr int SCR_MAXC = 3;   // objects a walker caches for the screen
constexpr unsigned SCR_LEVEL_SHIFT = 30;
constexpr unsigned SCR_LIST_MASK = (1u << SCR_LEVEL_SHIFT) - 1u;

DMW_SCR_HD float screen_box_width(float w, unsigned level) { return w * (1.0f - 0.5f * (float)level); }

// Distance margin of the screen (host).  Covers, with a factor of at least eight:
// the float32 rounding of positions and cached image centres (a few ulp of the
// cell side), of f (ulp of R^2 and |x - c|^2), of the reference's root near
// tangency (the square root of the rounding of its discriminant, ~2^-20 R as a
// distance), and of its unit direction (~2^-22 |s|).
inline double screen_margin(double cell, double r_max, double step) {
  return ldexp(cell + r_max + step, -16);
}

}  // namespace dmw
