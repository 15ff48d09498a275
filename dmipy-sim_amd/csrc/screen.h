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
// For an object of radius R centred at c, d2(x) = |x - c|^2 along the segment
// x(tau) = p + tau s, tau in [0, 1], is a convex parabola with leading
// coefficient |s|^2, so
//   max_tau d2 <= max(d2(p), d2(p + s))                    (convexity), and
//   min_tau d2 >= min(d2(p), d2(p + s)) - |s|^2 / 4         (vertex depth).
// The segment therefore stays at least a margin m outside the wall when both
// ends have d2 > a_out = (R + m)^2 + |s|^2 / 4, and inside it when both have
// d2 < a_in = (R - m)^2.  In either case the reference's root test finds no
// valid root for that object: its float32 roots and the screen's float32 values
// differ from exact geometry by far less than m, which the host sizes from the
// cell side, the largest radius and the step (screen_margin below; the bound is
// exercised against the oracle's float32 root test on adversarial configurations
// in tests/test_screen.py).
//
// The walk keeps, per cached object, the side its current position is clear on
// as a pair (rs, b) (screen_state): rs = -R inside, +R otherwise, and b chosen so
// that an end point passes exactly when fma(d2, rs, b) > 0, i.e. d2 > a_out
// (outside) or d2 < a_in (inside), and never when the position is within the
// margin band (b = -inf).  A sub-step whose end point passes for every object
// starts and ends clear on the same side of each wall, so the state stays valid
// and is not rewritten; only the walkers that do not pass run the reference's
// test, after which their state is set anew.
#pragma once

#include <math.h>

#if defined(__HIPCC__)
#define DMW_SCR_HD __host__ __device__ __forceinline__
#else
#define DMW_SCR_HD static inline
#endif

namespace dmw {

// |x - c|^2 in float32 (dim 2 ignores z).
DMW_SCR_HD float screen_d2(float px, float py, float pz, float cx, float cy, float cz, int dim) {
  const float ex = px - cx, ey = py - cy;
  float d2 = fmaf(ex, ex, ey * ey);
  if (dim == 3) {
    const float ez = pz - cz;
    d2 = fmaf(ez, ez, d2);
  }
  return d2;
}

// State of a walker at squared distance d2 from an object of radius R:
//   m2 = 2 m, c_out = m^2 + |s|max^2 / 4, c_in = m^2.
// The products R * a are rounded once; the test's fma is exact in sign, so the
// thresholds move by a relative 2^-24 at most, far inside the margin.
DMW_SCR_HD void screen_state(float d2, float R, float m2, float c_out, float c_in, float& rs, float& b) {
  const float a_out = fmaf(R, R, fmaf(R, m2, c_out));
  const float a_in = fmaf(R, R, fmaf(-R, m2, c_in));
  const bool out = d2 > a_out, in = d2 < a_in;
  rs = in ? -R : R;
  b = out ? -(R * a_out) : (in ? R * a_in : -INFINITY);
}

// True when an end point at squared distance d2 stays clear on the state's side.
DMW_SCR_HD bool screen_pass(float d2, float rs, float b) { return fmaf(d2, rs, b) > 0.0f; }

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
