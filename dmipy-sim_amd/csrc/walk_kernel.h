// walk_kernel.h -- the wavefront-resident Monte-Carlo walk for gfx950.
//
// One lane = one walker for the whole walk: position, phase, compensation,
// log-weight, compartment and crossing count live in VGPRs across all n_t
// waveform samples and n_sub sub-steps; nothing per step touches HBM.  The
// gradient G(t) is wave-uniform and is read with scalar loads; packed-substrate
// object tables and their periodic neighbour grid are staged in LDS.
//
// Semantics follow the fused engine of the reference:
//   sub-step       physics.py:296-314 (surface 337-355, plain 371-387, 402-412)
//   FreeDiffusion  geometries.py:40-41
//   Box1D          geometries.py:93-139
//   Sphere         geometries.py:197-317 (reflect*), 320-424 (permeate)
//   Cylinder       geometries.py:493-654 (reflect*), 656-801 (permeate)
//   PackedCyl.     geometries.py:1746-2016 (reflect*), 2018-2183 (permeate),
//                  2185-2214 (classify)
//   PackedSph.     geometries.py:2445-2631, 2633-2756, 2758-2771
//   Ellipsoid      geometries.py:1137-1425
//   Slab / Shell   geometries.py:3221-3300, 3303-3431
//   Myelinated     physics.py:420-679 (make_myelin_step_fn), geometries.py:847-1136
//   ensemble       core.py:426-429, stimulated echo 734-735 (host)
// with these deliberate differences, none of which changes a position or a wall
// decision:
//   * the bounce loop stops at the first iteration without a hit (the
//     reference's fixed 10 iterations are no-ops after that);
//   * packed substrates test only the objects listed in the walker's periodic
//     grid cell (the others cannot be reached within one step; first-index
//     argmin order is preserved by the ascending cell lists);
//   * packed positions are an integer cell index plus a float32 offset, folded
//     once per waveform sample;
//   * the phase is advanced once per waveform sample with the sum of the n_sub
//     sub-step positions (G is constant across the sub-steps of a sample,
//     physics.py:287-289), optionally Kahan-compensated; on the matrix cores the
//     walker operand is that sum less n_sub times the start point, both operands
//     split into float16 hi + lo parts (relative residual below 2^-21), and the
//     start point's phase (start . sum_t G) is added back at the end;
//   * the ensemble sum is float64, reduced in a fixed order.
// All float32 wall arithmetic is written operation by operation in the
// reference's order and compiled without contraction.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>
#include <utility>

#include "dmw_rng.h"
#include "rn_arith.h"
#include "screen.h"

#pragma clang fp contract(off)

namespace dmw {

enum GeomKind {
  G_FREE = 0, G_BOX1D = 1, G_CYL = 2, G_SPH = 3, G_PCYL = 4, G_PSPH = 5,
  G_ELL = 6,    // Ellipsoid: axes = 1/a^2, 1/b^2, 1/c^2; aux = 1 - NUDGE/min(a,b,c)
  G_SLAB = 7,   // PermeableSlab1D: size = L, membrane at L/2
  G_SHELL = 8,  // PermeableShell, spherical: size = R_in, aux = R_out
  G_SHELLC = 9, // PermeableShell, cylindrical (axis z in the native frame)
  G_MYEL = 10,  // MyelinatedCylinder: size = R_in, aux = R_out (axis z in the native frame)
  G_PMYEL = 11  // PackedMyelinatedCylinders: objs rows (cx, cy, R_in, R_out), steps, 2 kappa/D
};
enum WallKind { W_REFLECT = 0, W_SURFACE = 1, W_PERMEATE = 2 };

typedef __attribute__((address_space(4))) const float cfloat;

// Compile-time loop: f(std::integral_constant<int, i>) for i in [0, N), so array
// indices stay constants and per-walker arrays stay in registers.
template <class F, int... I>
__device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, I...>) {
  (f(std::integral_constant<int, I>{}), ...);
}
template <int N, class F>
__device__ __forceinline__ void static_for(F&& f) {
  static_for_impl(f, std::make_integer_sequence<int, N>{});
}


constexpr int kBlock = 256;
constexpr int kWave = 64;

struct WalkArgs {
  // geometry
  float size, cell, inv_cell, eps, nudge, rhoD, kappaD;
  float inv_size;            // RN(1 / size): reciprocal for the known-divisor quotient
  float size_sq_cut;         // least float s with sqrtf(s) >= size: sqrt_rn(s) >= size <=> s >= cut
  float size_sq_cut_gt;      // least float s with sqrtf(s) > size
  int n_obj;
  const float4* objs;        // device (cx, cy, cz, R), native frame
  int grid_n;                // packed: grid cells per axis
  float grid_h, grid_inv_h;  // packed: grid cell side
  const int* cell_start;     // device (grid_n^d + 1)
  const int* cell_items;     // device
  int n_items;
  int exact_minimage;        // 1: min-image folds by true division
  // screened packed permeation: cells expanded by scr_w on every side; scr_cache =
  // 1 caches the cell's objects for the wall screen (screen.h) with margin m:
  // scr_m2 = 2 m, scr_cout = m^2 + step^2 / 4, scr_cin = m^2
  int scr_cache;
  float scr_w, scr_m2, scr_cout, scr_cin;
  // step
  int n_sub;
  float step_len, dt_sub, inv_T2, inv_T1;
  int has_t2, has_t1;
  // waveform
  int n_t, n_meas, n_meas_pad;
  const float* Gs;           // device (n_t, n_meas_pad, 3), scaled by gamma*dt_sub
  // Matrix-core phase operands: Gs * g_scale split into float16 hi and lo parts,
  // in 16-byte per-lane chunks (n_t, 2, n_meas_pad / 32, 16) x (2 x (x, y, z, 0));
  // Gsum = sums over t of Gs between emission points, (n_echo + 1, 3, n_meas_pad).
  // The walker operand is its displacement sum from the start point times
  // p_scale (both scales powers of two, chosen so neither operand exceeds 2^14).
  const uint2* GsH;
  const float* Gsum;
  float p_scale, inv_pg_scale;  // inv_pg_scale = 1 / (p_scale * g_scale)
  const float* chi;          // device (n_t) or null
  int n_echo;
  const int* echo_idx;       // device (n_echo)
  // walkers: this launch walks walkers [first_walker, first_walker + n_walkers) of
  // a walk; the global index keys the Philox streams and the in-kernel start rules
  long long n_walkers, first_walker;
  uint32_t seed_lo, seed_hi, instance;
  float ax, ay, az, aux;     // Ellipsoid 1/semi-axis^2 and clamp factor; shell R_out
  // MyelinatedCylinder, per compartment (intra, sheath, extra): step length, dt/T2,
  // spin weight; 2 kappa/D_leaving for (inner from 0, inner from 1, outer from 1,
  // outer from 2); walkers [0, m_n0) start intra, [m_n0, m_n01) in the sheath.
  float m_step[3], m_q2[3], m_wf[3], m_k2[4];
  long long m_n0, m_n01;
  // PackedMyelinatedCylinders: dt_sub/T2 per class (intra, myelin, extra), wall
  // relaxivity / D_ref, myelin spin weight, start boundaries (2 n_obj + 1, device).
  float pm_q2[3], pm_rhoD, pm_wmy;
  const long long* pm_start;
  const float* r0;           // device (3, N) or null
  const float* inj_n;        // device (S, 3, N) or null, S = n_t * n_sub
  const float* inj_u;        // device (S, N) or null
  // outputs
  float4* park;              // parked phase: (grid.y, waves, MB/4, 2, 64) float4, or null
  // Samples [t_begin, t_end) of the walk.  A launch with t_begin > 0 resumes the
  // walkers from state_in; one with t_end < n_t writes their state to state_out
  // instead of the ensemble.  State: three float4 rows of N, (x, y, z, log w),
  // (cell x, cell y, cell z, crossings) as int bits, (unwrapped x, y, z, 0).
  int t_begin, t_end;
  const float4* state_in;
  float4* state_out;
  double* partials;          // (n_waves, n_meas_pad, 2)
  double* echo_partials;     // (n_waves, n_echo, n_meas_pad, 2)
  float* final_pos;          // (3, N)
  int* comp_origin;          // (N)
  int* comp_final;           // (N)
  int* comp_full;            // (n_t, N)
  float* pos_full;           // (n_t, 3, N)
  uint32_t* n_cross;         // (N)
  float* log_w;              // (N)
  float* phi;                // (n_meas_pad, N)
  unsigned long long* counters;  // [2]: wall events, candidates tested
  // trajectory channels of the walk-once producer (float16 bits, SoA)
  uint16_t* traj_h;          // (n_t, 3, N): unfolded position after every sample
  uint16_t* dlog_h;          // (n_t, N): wall log-weight summed over the sample's sub-steps
  uint16_t* comp_h;          // (n_t, N): fraction of the sample's sub-steps outside (permeable Cylinder/Sphere)
  int traj_rotate;           // 1: store traj_R . position (lab frame of a rotated geometry)
  float traj_R[9];           // row-major lab-from-native rotation
};

// ---------------------------------------------------------------------------
// Walker position: offset (float32) + periodic cell index (packed substrates).
// ---------------------------------------------------------------------------
struct Pos {
  float x, y, z;
  int cx, cy, cz;
};

// Source of the Powles uniform: injected, or a Philox draw made only when a
// wall is actually hit (the reference draws one every sub-step and uses it only
// on a hit, geometries.py:2122-2125).
struct PowlesSrc {
  uint32_t walker, instance, gstep, seed_lo, seed_hi;
  float injected;
  bool inject;
  __device__ __forceinline__ float get() const {
    if (inject) return injected;
    u32x4 r = philox4x32_10({walker, instance, gstep, TAG_POWLES}, seed_lo, seed_hi);
    return u01(r.x);
  }
};

// floor(q / L + 0.5) as the reference computes it (true division).  For
// objects listed in a walker's grid cell the quotient is never near a rounding
// boundary (checked on the host), so the reciprocal form gives the same integer.
__device__ __forceinline__ float img_shift(float q, const WalkArgs& a) {
  if (a.exact_minimage) return floorf(q / a.cell + 0.5f);
  return floorf(q * a.inv_cell + 0.5f);
}

// Exact fold of a position into the primary cell: x - L*floor(x/L + 0.5)
// (geometries.py:1847).  Uses true division only when the quotient is close to
// a rounding boundary.
__device__ __forceinline__ float fold_shift(float x, const WalkArgs& a) {
  const float app = x * a.inv_cell + 0.5f;
  const float k = floorf(app);
  if (fabsf(app - rintf(app)) < 1e-5f) return floorf(x / a.cell + 0.5f);
  return k;
}

__device__ __forceinline__ int grid_index(float v, const WalkArgs& a) {
  const float f = v - a.cell * floorf(v * a.inv_cell + 0.5f);
  int i = (int)floorf((f + 0.5f * a.cell) * a.grid_inv_h);
  i = i < 0 ? 0 : i;
  return i >= a.grid_n ? a.grid_n - 1 : i;
}

// ---------------------------------------------------------------------------
// Box1D (geometries.py:93-139)
// ---------------------------------------------------------------------------
template <int WALL>
__device__ __forceinline__ float box_step(Pos& p, float sx, float sy, float sz, const WalkArgs& a,
                                          uint32_t& hits) {
  const float L = a.size;
  const float L2 = 2.0f * L;
  const float xr = p.x + sx;
  float dlog = 0.0f;
  if (WALL == W_SURFACE) {
    const float dperp = fmaxf(xr - L, 0.0f) + fmaxf(-xr, 0.0f);
    if (dperp > 0.0f) { dlog = (-2.0f * a.rhoD) * dperp; hits++; }
  }
  float m = fmodf(xr, L2);
  if (m != 0.0f && ((m < 0.0f) != (L2 < 0.0f))) m = m + L2;
  if (m > L) m = L2 - m;
  p.x = m;
  p.y = p.y + sy;
  p.z = p.z + sz;
  return dlog;
}

// ---------------------------------------------------------------------------
// Bounce-loop shortcut after a reflection.  With the walker inside (c = |p|^2 - R^2
// < 0, as computed by the iteration) and heading inward (dp = u . p <= -2^-60, so
// dp^2 does not underflow), the iteration's float32 quantities satisfy
// disc >= fl(dp^2), sqrt_rn(disc) >= |dp| (1 - 2^-23) and d = -dp + sq >= 2 |dp|
// (1 - 2^-22).  A remaining path rem <= fl(2 |dp| (1 - 2^-20)) is then below d, so
// the iteration would find no hit and leave every value unchanged: it can be skipped
// (its square root is the saving).  Only used after a hit, where it decides for
// almost every lane that reflected.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool no_second_hit(float dp, float c, float rem) {
  return dp <= -0x1p-60f && c < 0.0f && rem <= (-2.0f * dp) * 0x1.ffffep-1f;
}

// ---------------------------------------------------------------------------
// Cylinder, interior reflection (geometries.py:516-654).  Returns sum d_perp.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float cyl_reflect(float& x, float& y, float sx, float sy, const WalkArgs& a,
                                             uint32_t& hits) {
  const float R = a.size, EPS = a.eps, NUDGE = a.nudge;
  const float RR = R * R;
  const float sl = sqrt_rn(sx * sx + sy * sy);
  float dx = 0.0f, dy = 0.0f;
  if (sl > 0.0f) { dx = div_rn(sx, sl); dy = div_rn(sy, sl); }
  float rem = sl, dsum = 0.0f;
  for (int it = 0; it < 10; ++it) {
    const float dp = dx * x + dy * y;
    const float c = (x * x + y * y) - RR;
    if (it > 0 && no_second_hit(dp, c, rem)) break;
    const float disc = fmaxf(dp * dp - c, 0.0f);
    const float sq = sqrt_rn(disc);
    const float d = -dp + sq;
    if (!((d > EPS) && (d < rem) && (sl > 0.0f))) break;
    const float hx = x + d * dx, hy = y + d * dy;
    const float nx = div_rn_r(hx, R, a.inv_size), ny = div_rn_r(hy, R, a.inv_size);
    const float tdd = 2.0f * (dx * nx + dy * ny);
    float rx = dx - tdd * nx, ry = dy - tdd * ny;
    const float rn = sqrt_rn(rx * rx + ry * ry);
    rx = div_rn(rx, rn); ry = div_rn(ry, rn);
    dsum = dsum + (rem - d) * (div_rn_r(sq, R, a.inv_size));
    x = hx - NUDGE * nx;
    y = hy - NUDGE * ny;
    dx = rx; dy = ry;
    rem = (rem - d) - NUDGE;
    hits++;
  }
  const float rm = fmaxf(rem, 0.0f);
  float fx = x + dx * rm, fy = y + dy * rm;
  // fn >= R decided on the squared norm (exact: sqrt_rn is monotone, size_sq_cut is the
  // least square whose root reaches R); the root itself only inside the rare clamp
  const float ss = fx * fx + fy * fy;
  if (ss >= a.size_sq_cut) {
    const float fn = sqrt_rn(ss);
    const float s = R - NUDGE;
    fx = div_rn(fx * s, fn);
    fy = div_rn(fy * s, fn);
  }
  x = fx; y = fy;
  return dsum;
}

// Cylinder, Powles permeation (geometries.py:656-801).
__device__ __forceinline__ float cyl_permeate(float& x, float& y, float sx, float sy, const WalkArgs& a,
                                              const PowlesSrc& ps, uint32_t& hits, uint32_t& ncross) {
  const float R = a.size, EPS = a.eps, NUDGE = a.nudge;
  const float RR = R * R;
  const float sl = sqrt_rn(sx * sx + sy * sy);
  float dx = 0.0f, dy = 0.0f;
  if (sl > 0.0f) { dx = div_rn(sx, sl); dy = div_rn(sy, sl); }
  const float dp = dx * x + dy * y;
  const float rr = x * x + y * y;
  const float disc = dp * dp - (rr - RR);
  const float sq = sqrt_rn(fmaxf(disc, 0.0f));
  const bool inside = rr < RR;
  const float t_hit = inside ? (-dp + sq) : (-dp - sq);
  const bool any_hit = (disc > 0.0f) && (t_hit > EPS) && (t_hit < sl) && (sl > 0.0f);
  if (!any_hit) { x = x + sx; y = y + sy; return 0.0f; }
  hits++;
  const float hx = x + t_hit * dx, hy = y + t_hit * dy;
  const float hl = fmaxf(sqrt_rn(hx * hx + hy * hy), 1e-30f);
  const float nx = div_rn(hx, hl), ny = div_rn(hy, hl);
  const float remaining = sl - t_hit;
  const float dperp = remaining * (div_rn_r(sq, R, a.inv_size));
  const float p = fminf(1.0f, (2.0f * a.kappaD) * dperp);
  if (ps.get() < p) {
    x = x + sx; y = y + sy; ncross++;
    return 0.0f;
  }
  const float tdd = 2.0f * (dx * nx + dy * ny);
  float rx = dx - tdd * nx, ry = dy - tdd * ny;
  const float rn = sqrt_rn(rx * rx + ry * ry);
  rx = div_rn(rx, rn); ry = div_rn(ry, rn);
  const float sxh = R * nx, syh = R * ny;
  const float qx = inside ? sxh - NUDGE * nx : sxh + NUDGE * nx;
  const float qy = inside ? syh - NUDGE * ny : syh + NUDGE * ny;
  const float go = fmaxf(remaining - NUDGE, 0.0f);
  float fx = qx + rx * go, fy = qy + ry * go;
  // fnorm >= R and fnorm <= R decided on the squared norm (as in cyl_reflect; fnorm > R
  // exactly when the square reaches size_sq_cut_gt), the root only for the clamp
  const float ss = fx * fx + fy * fy;
  const float target = inside ? R - NUDGE : R + NUDGE;
  const bool wrong = inside ? (ss >= a.size_sq_cut) : (ss < a.size_sq_cut_gt);
  if (wrong) {
    const float fsafe = fmaxf(sqrt_rn(ss), 1e-30f);
    fx = div_rn(fx * target, fsafe);
    fy = div_rn(fy * target, fsafe);
  }
  x = fx; y = fy;
  return (-2.0f * a.rhoD) * dperp;
}

// ---------------------------------------------------------------------------
// Sphere, interior reflection (geometries.py:197-317).  Returns sum d_perp.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float sph_reflect(Pos& p, float sx, float sy, float sz, const WalkArgs& a,
                                             uint32_t& hits) {
  const float R = a.size, EPS = a.eps, NUDGE = a.nudge;
  const float RR = R * R;
  const float sl = sqrt_rn((sx * sx + sy * sy) + sz * sz);
  float dx = div_rn(sx, sl), dy = div_rn(sy, sl), dz = div_rn(sz, sl);
  float x = p.x, y = p.y, z = p.z;
  float rem = sl, dsum = 0.0f;
  for (int it = 0; it < 10; ++it) {
    const float dp = (dx * x + dy * y) + dz * z;
    const float c = ((x * x + y * y) + z * z) - RR;
    if (it > 0 && no_second_hit(dp, c, rem)) break;
    const float disc = fmaxf(dp * dp - c, 0.0f);
    const float sq = sqrt_rn(disc);
    const float d = -dp + sq;
    if (!((d > EPS) && (d < rem))) break;
    const float hx = x + d * dx, hy = y + d * dy, hz = z + d * dz;
    const float nx = div_rn_r(hx, R, a.inv_size), ny = div_rn_r(hy, R, a.inv_size);
    const float nz = div_rn_r(hz, R, a.inv_size);
    const float tdd = 2.0f * ((dx * nx + dy * ny) + dz * nz);
    float rx = dx - tdd * nx, ry = dy - tdd * ny, rz = dz - tdd * nz;
    const float rn = sqrt_rn((rx * rx + ry * ry) + rz * rz);
    rx = div_rn(rx, rn); ry = div_rn(ry, rn); rz = div_rn(rz, rn);
    dsum = dsum + (rem - d) * (div_rn_r(sq, R, a.inv_size));
    x = hx - NUDGE * nx; y = hy - NUDGE * ny; z = hz - NUDGE * nz;
    dx = rx; dy = ry; dz = rz;
    rem = (rem - d) - NUDGE;
    hits++;
  }
  const float rm = fmaxf(rem, 0.0f);
  float fx = x + dx * rm, fy = y + dy * rm, fz = z + dz * rm;
  const float ss = (fx * fx + fy * fy) + fz * fz;
  if (ss >= a.size_sq_cut) {  // fn >= R on the squared norm, as in cyl_reflect
    const float fn = sqrt_rn(ss);
    const float s = R - NUDGE;
    fx = div_rn(fx * s, fn); fy = div_rn(fy * s, fn); fz = div_rn(fz * s, fn);
  }
  p.x = fx; p.y = fy; p.z = fz;
  return dsum;
}

// Sphere, Powles permeation (geometries.py:320-424): no snap, no side clamp.
__device__ __forceinline__ float sph_permeate(Pos& p, float sx, float sy, float sz, const WalkArgs& a,
                                              const PowlesSrc& ps, uint32_t& hits, uint32_t& ncross) {
  const float R = a.size, EPS = a.eps, NUDGE = a.nudge;
  const float RR = R * R;
  const float sl = sqrt_rn((sx * sx + sy * sy) + sz * sz);
  const float dx = div_rn(sx, sl), dy = div_rn(sy, sl), dz = div_rn(sz, sl);
  const float x = p.x, y = p.y, z = p.z;
  const float dp = (dx * x + dy * y) + dz * z;
  const float rr = (x * x + y * y) + z * z;
  const float disc = dp * dp - (rr - RR);
  const float sq = sqrt_rn(fmaxf(disc, 0.0f));
  const bool inside = rr < RR;
  const float t_hit = inside ? (-dp + sq) : (-dp - sq);
  const bool any_hit = (disc > 0.0f) && (t_hit > EPS) && (t_hit < sl) && (sl > 0.0f);
  if (!any_hit) { p.x = x + sx; p.y = y + sy; p.z = z + sz; return 0.0f; }
  hits++;
  const float hx = x + t_hit * dx, hy = y + t_hit * dy, hz = z + t_hit * dz;
  const float nx = div_rn_r(hx, R, a.inv_size), ny = div_rn_r(hy, R, a.inv_size);
  const float nz = div_rn_r(hz, R, a.inv_size);
  const float remaining = sl - t_hit;
  const float dperp = remaining * (div_rn_r(sq, R, a.inv_size));
  const float pr = fminf(1.0f, (2.0f * a.kappaD) * dperp);
  if (ps.get() < pr) {
    p.x = x + sx; p.y = y + sy; p.z = z + sz; ncross++;
    return 0.0f;
  }
  const float tdd = 2.0f * ((dx * nx + dy * ny) + dz * nz);
  float rx = dx - tdd * nx, ry = dy - tdd * ny, rz = dz - tdd * nz;
  const float rn = sqrt_rn((rx * rx + ry * ry) + rz * rz);
  rx = div_rn(rx, rn); ry = div_rn(ry, rn); rz = div_rn(rz, rn);
  const float qx = inside ? hx - NUDGE * nx : hx + NUDGE * nx;
  const float qy = inside ? hy - NUDGE * ny : hy + NUDGE * ny;
  const float qz = inside ? hz - NUDGE * nz : hz + NUDGE * nz;
  const float go = fmaxf(remaining - NUDGE, 0.0f);
  p.x = qx + rx * go; p.y = qy + ry * go; p.z = qz + rz * go;
  return (-2.0f * a.rhoD) * dperp;
}

// ---------------------------------------------------------------------------
// Ellipsoid with semi-axes (a, b, c) along x, y, z (geometries.py:1137-1425).
// The quadratic A t^2 + 2 B t + C of the ray against x.Dx = 1, D = diag(1/a^2 ...).
// ---------------------------------------------------------------------------
__device__ __forceinline__ float ell_form(float ux, float uy, float uz, float vx, float vy, float vz,
                                          const WalkArgs& a) {
  return ((ux * a.ax) * vx + (uy * a.ay) * vy) + (uz * a.az) * vz;
}

// Reflection with the 10-iteration bounce scan and the normalised-coordinate
// clamp; returns sum d_perp = (remaining - d) cos(alpha) (geometries.py:1187-1300).
__device__ __forceinline__ float ell_reflect(Pos& p, float sx, float sy, float sz, const WalkArgs& a,
                                             uint32_t& hits) {
  const float EPS = a.eps, NUDGE = a.nudge;
  const float sl = sqrt_rn((sx * sx + sy * sy) + sz * sz);
  float dx = div_rn(sx, sl), dy = div_rn(sy, sl), dz = div_rn(sz, sl);
  float x = p.x, y = p.y, z = p.z, rem = sl, dsum = 0.0f;
  for (int it = 0; it < 10; ++it) {
    const float A = ell_form(dx, dy, dz, dx, dy, dz, a);
    const float B = ell_form(x, y, z, dx, dy, dz, a);
    const float C = ell_form(x, y, z, x, y, z, a) - 1.0f;
    const float disc = fmaxf(B * B - A * C, 0.0f);
    const float d = div_rn(-B + sqrt_rn(disc), A);
    if (!((d > EPS) && (d < rem))) break;
    const float hx = x + d * dx, hy = y + d * dy, hz = z + d * dz;
    const float mx = hx * a.ax, my = hy * a.ay, mz = hz * a.az;
    const float mn = sqrt_rn((mx * mx + my * my) + mz * mz);
    const float nx = div_rn(mx, mn), ny = div_rn(my, mn), nz = div_rn(mz, mn);
    const float cosa = (dx * nx + dy * ny) + dz * nz;
    const float tdd = 2.0f * cosa;
    float rx = dx - tdd * nx, ry = dy - tdd * ny, rz = dz - tdd * nz;
    const float rn = sqrt_rn((rx * rx + ry * ry) + rz * rz);
    rx = div_rn(rx, rn); ry = div_rn(ry, rn); rz = div_rn(rz, rn);
    dsum = dsum + (rem - d) * cosa;
    x = hx - NUDGE * nx; y = hy - NUDGE * ny; z = hz - NUDGE * nz;
    dx = rx; dy = ry; dz = rz;
    rem = (rem - d) - NUDGE;
    hits++;
  }
  const float rm = fmaxf(rem, 0.0f);
  float fx = x + dx * rm, fy = y + dy * rm, fz = z + dz * rm;
  const float qx = fx * sqrt_rn(a.ax), qy = fy * sqrt_rn(a.ay), qz = fz * sqrt_rn(a.az);
  const float q = sqrt_rn((qx * qx + qy * qy) + qz * qz);
  if (q >= 1.0f) { fx = div_rn(fx * a.aux, q); fy = div_rn(fy * a.aux, q); fz = div_rn(fz * a.aux, q); }
  p.x = fx; p.y = fy; p.z = fz;
  return dsum;
}

// Powles permeation, single event per step (geometries.py:1302-1410).
__device__ __forceinline__ float ell_permeate(Pos& p, float sx, float sy, float sz, const WalkArgs& a,
                                              const PowlesSrc& ps, uint32_t& hits, uint32_t& ncross) {
  const float EPS = a.eps, NUDGE = a.nudge;
  const float sl = sqrt_rn((sx * sx + sy * sy) + sz * sz);
  const float dx = div_rn(sx, sl), dy = div_rn(sy, sl), dz = div_rn(sz, sl);
  const float x = p.x, y = p.y, z = p.z;
  const float A = ell_form(dx, dy, dz, dx, dy, dz, a);
  const float B = ell_form(x, y, z, dx, dy, dz, a);
  const float C = ell_form(x, y, z, x, y, z, a) - 1.0f;
  const float disc_raw = B * B - A * C;
  const float sq = sqrt_rn(fmaxf(disc_raw, 0.0f));
  const bool inside = C < 0.0f;
  const float t_hit = inside ? div_rn(-B + sq, A) : div_rn(-B - sq, A);
  const bool any_hit = (disc_raw > 0.0f) && (t_hit > EPS) && (t_hit < sl) && (sl > 0.0f);
  if (!any_hit) { p.x = x + sx; p.y = y + sy; p.z = z + sz; return 0.0f; }
  hits++;
  const float hx = x + t_hit * dx, hy = y + t_hit * dy, hz = z + t_hit * dz;
  const float mx = hx * a.ax, my = hy * a.ay, mz = hz * a.az;
  const float mn = sqrt_rn((mx * mx + my * my) + mz * mz);
  const float nx = div_rn(mx, mn), ny = div_rn(my, mn), nz = div_rn(mz, mn);
  const float remaining = sl - t_hit;
  const float dn = (dx * nx + dy * ny) + dz * nz;
  const float dperp = remaining * fabsf(dn);
  const float pr = fminf(1.0f, (2.0f * a.kappaD) * dperp);
  if (ps.get() < pr) {
    p.x = x + sx; p.y = y + sy; p.z = z + sz; ncross++;
    return 0.0f;
  }
  const float tdd = 2.0f * dn;
  float rx = dx - tdd * nx, ry = dy - tdd * ny, rz = dz - tdd * nz;
  const float rn = sqrt_rn((rx * rx + ry * ry) + rz * rz);
  rx = div_rn(rx, rn); ry = div_rn(ry, rn); rz = div_rn(rz, rn);
  const float ux = inside ? -nx : nx, uy = inside ? -ny : ny, uz = inside ? -nz : nz;
  const float qx = hx + NUDGE * ux, qy = hy + NUDGE * uy, qz = hz + NUDGE * uz;
  const float go = fmaxf(remaining - NUDGE, 0.0f);
  p.x = qx + rx * go; p.y = qy + ry * go; p.z = qz + rz * go;
  return (-2.0f * a.rhoD) * dperp;
}

// ---------------------------------------------------------------------------
// PermeableSlab1D: membrane at x = L/2, reflecting walls at 0 and L folded by
// the modular mirror (geometries.py:3221-3300).
// ---------------------------------------------------------------------------
__device__ __forceinline__ float slab_fold(float x, float L) {
  const float L2 = 2.0f * L;
  float m = fmodf(x, L2);
  if (m != 0.0f && ((m < 0.0f) != (L2 < 0.0f))) m = m + L2;
  return m > L ? L2 - m : m;
}

template <int WALL>
__device__ __forceinline__ float slab_step(Pos& p, float sx, float sy, float sz, const WalkArgs& a,
                                           const PowlesSrc& ps, uint32_t& hits, uint32_t& ncross) {
  const float L = a.size, xm = 0.5f * a.size;
  const float x = p.x, xn = x + sx;
  const bool crossed = (x - xm) * (xn - xm) < 0.0f;
  float dlog = 0.0f, x1 = xn;
  if (crossed) {
    hits++;
    const float dperp = fabsf(xn - xm);
    bool transmit = false;
    if (WALL == W_PERMEATE) {
      const float pr = fminf(1.0f, (2.0f * a.kappaD) * dperp);
      transmit = ps.get() < pr;
    }
    if (transmit) {
      ncross++;
    } else {
      x1 = 2.0f * xm - xn;
      if (WALL != W_REFLECT) dlog = (-2.0f * a.rhoD) * dperp;
    }
  }
  p.x = slab_fold(x1, L);
  p.y = p.y + sy;
  p.z = p.z + sz;
  return dlog;
}

// ---------------------------------------------------------------------------
// PermeableShell: permeable membrane at R_in inside a reflecting wall at R_out;
// spherical (radial = r) or cylindrical (radial = (x, y, 0) in the native frame)
// (geometries.py:3303-3431).  The reflecting form is the same routine with
// kappa = rho = 0 (geometries.py:3428-3431).
// ---------------------------------------------------------------------------
template <bool CYLK>
__device__ __forceinline__ float shell_norm(float x, float y, float z) {
  return CYLK ? sqrt_rn((x * x + y * y) + 0.0f * 0.0f) : sqrt_rn((x * x + y * y) + z * z);
}

template <bool CYLK>
__device__ __forceinline__ float shell_first_t(float b, float r2, float dd2, float Rk, float EPS, float sl) {
  const float c = r2 - Rk * Rk;
  const float disc = b * b - dd2 * c;
  const float sq = sqrt_rn(fmaxf(disc, 0.0f));
  float t1 = div_rn(-b - sq, dd2), t2 = div_rn(-b + sq, dd2);
  t1 = (disc > 0.0f && t1 > EPS && t1 < sl) ? t1 : 1e30f;
  t2 = (disc > 0.0f && t2 > EPS && t2 < sl) ? t2 : 1e30f;
  return fminf(t1, t2);
}

template <bool CYLK, int WALL>
__device__ __forceinline__ float shell_step(Pos& p, float sx, float sy, float sz, const WalkArgs& a,
                                            const PowlesSrc& ps, uint32_t& hits, uint32_t& ncross) {
  const float Rin = a.size, Rout = a.aux, EPS = a.eps, NUDGE = a.nudge;
  const float sl = sqrt_rn((sx * sx + sy * sy) + sz * sz);
  const float den = fmaxf(sl, EPS);
  const float dx = div_rn(sx, den), dy = div_rn(sy, den), dz = div_rn(sz, den);
  const float x = p.x, y = p.y, z = p.z;
  const float rz = CYLK ? 0.0f : z, dzr = CYLK ? 0.0f : dz;       // radial components
  const float b = (dx * x + dy * y) + dzr * rz;
  const float r2 = (x * x + y * y) + rz * rz;
  const float dd2 = fmaxf((dx * dx + dy * dy) + dzr * dzr, EPS);
  const float t_in = shell_first_t<CYLK>(b, r2, dd2, Rin, EPS, sl);
  const float t_out = shell_first_t<CYLK>(b, r2, dd2, Rout, EPS, sl);
  const bool hit_in = t_in < t_out;
  const float t_hit = fminf(t_in, t_out);
  const bool any_hit = t_hit < 1e29f;
  float ox = x + sx, oy = y + sy, oz = z + sz;
  float dlog = 0.0f;
  if (any_hit) {
    hits++;
    const float hx = x + t_hit * dx, hy = y + t_hit * dy, hz = z + t_hit * dz;
    const float hzr = CYLK ? 0.0f : hz;
    const float hn = fmaxf(sqrt_rn((hx * hx + hy * hy) + hzr * hzr), EPS);
    const float nx = div_rn(hx, hn), ny = div_rn(hy, hn), nz = div_rn(hzr, hn);
    const float remaining = sl - t_hit;
    const float dn = (dx * nx + dy * ny) + dz * nz;
    const float dperp = remaining * fabsf(dn);
    bool transmit = false;
    if (WALL == W_PERMEATE && hit_in) {
      const float pr = fminf(1.0f, (2.0f * a.kappaD) * dperp);
      transmit = ps.get() < pr;
    }
    if (transmit) {
      ncross++;
    } else {
      const float tdd = 2.0f * dn;
      const float qx = dx - tdd * nx, qy = dy - tdd * ny, qz = dz - tdd * nz;
      const float Rk = hit_in ? Rin : Rout;
      const float rmag = shell_norm<CYLK>(x, y, z);
      const float side = rmag < Rk ? -1.0f : 1.0f;
      const float sn = side * NUDGE;
      const float go = fmaxf(remaining - NUDGE, 0.0f);
      ox = (hx + sn * nx) + qx * go;
      oy = (hy + sn * ny) + qy * go;
      oz = (hz + sn * nz) + qz * go;
      if (WALL == W_PERMEATE && hit_in) dlog = (-2.0f * a.rhoD) * dperp;
    }
  }
  // never outside the reflecting wall (mirror the overshoot back)
  const float ozr = CYLK ? 0.0f : oz;
  const float rm = sqrt_rn((ox * ox + oy * oy) + ozr * ozr);
  const float over = rm - Rout;
  if (over > 0.0f) {
    const float rs = fmaxf(rm, EPS), t2 = 2.0f * over;
    ox = ox - t2 * (div_rn(ox, rs));
    oy = oy - t2 * (div_rn(oy, rs));
    oz = oz - t2 * (div_rn(ozr, rs));
  }
  p.x = ox; p.y = oy; p.z = oz;
  return dlog;
}

// ---------------------------------------------------------------------------
// MyelinatedCylinder (physics.py:513-668): compartments 0 = intra-axonal
// (r < R_in), 1 = sheath, 2 = extra-axonal (r >= R_out) in the native frame,
// carried in Pos::cx (the start compartment, which selects the spin weight,
// in Pos::cy).  One step per waveform sample with the compartment's own step
// length; a step that ends across a membrane of its compartment transmits
// with p = min(1, 2 kappa/D_leaving |r_new - R|), else it is mirrored to
// 2R - r_new radially; then the wrong-side clamps to R -+ NUDGE.  At most one
// membrane can be crossed per step, so one uniform decides it.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float pick3(int c, float v0, float v1, float v2) {
  return c == 0 ? v0 : (c == 1 ? v1 : v2);
}

__device__ __forceinline__ int myel_classify(float x, float y, const WalkArgs& a) {
  const float rr = x * x + y * y;
  return rr < a.size * a.size ? 0 : (rr < a.aux * a.aux ? 1 : 2);
}

// Rescale (x, y) to radius Rt when the wrong-side test holds (geometries'
// final_xy_hat * (R -+ NUDGE), physics.py:625-653).
__device__ __forceinline__ void myel_clamp(float& x, float& y, float& n, bool wrong, float Rt) {
  if (wrong) {
    const float s = fmaxf(n, 1e-20f);
    x = (div_rn(x, s)) * Rt;
    y = (div_rn(y, s)) * Rt;
    n = sqrt_rn(x * x + y * y);
  }
}

__device__ __forceinline__ void myel_step(Pos& p, float sx, float sy, float sz, const WalkArgs& a,
                                          const PowlesSrc& ps, uint32_t& hits, uint32_t& ncross) {
  const float Rin = a.size, Rout = a.aux, NUDGE = a.nudge;
  const int c = p.cx;
  float nx = p.x + sx, ny = p.y + sy;
  const float nz = p.z + sz;
  const float n = sqrt_rn(nx * nx + ny * ny);
  const bool in_out = (c == 0) && (n >= Rin), in_in = (c == 1) && (n < Rin);
  const bool out_out = (c == 1) && (n >= Rout), out_in = (c == 2) && (n < Rout);
  int nc = c;
  float nn = n;
  if (in_out || in_in || out_out || out_in) {
    hits++;
    const bool inner = in_out || in_in;
    const float Rk = inner ? Rin : Rout;
    const float k2 = inner ? (c == 0 ? a.m_k2[0] : a.m_k2[1]) : (c == 1 ? a.m_k2[2] : a.m_k2[3]);
    const float pr = fminf(1.0f, k2 * fabsf(n - Rk));
    if (ps.get() < pr) {
      ncross++;
      nc = inner ? (in_out ? 1 : 0) : (out_out ? 2 : 1);
    } else {
      const float s = fmaxf(n, 1e-20f);
      const float hx = div_rn(nx, s), hy = div_rn(ny, s);
      const float rr = 2.0f * Rk - n;
      nx = hx * rr;
      ny = hy * rr;
      nn = sqrt_rn(nx * nx + ny * ny);
    }
  }
  myel_clamp(nx, ny, nn, (nc == 0) && (nn >= Rin), Rin - NUDGE);
  myel_clamp(nx, ny, nn, (nc == 1) && (nn < Rin), Rin + NUDGE);
  myel_clamp(nx, ny, nn, (nc == 1) && (nn >= Rout), Rout - NUDGE);
  myel_clamp(nx, ny, nn, (nc == 2) && (nn < Rout), Rout + NUDGE);
  p.x = nx; p.y = ny; p.z = nz;
  p.cx = nc;
}

// ---------------------------------------------------------------------------
// Packed cylinders (2-D cell) and packed spheres (3-D cell).
// ---------------------------------------------------------------------------
struct Tables {
  const float4* objs;
  const int* start;
  const int* items;
};

template <int DIM>
__device__ __forceinline__ int cell_of(const Pos& p, const WalkArgs& a) {
  const int ix = grid_index(p.x, a), iy = grid_index(p.y, a);
  if (DIM == 2) return iy * a.grid_n + ix;
  const int iz = grid_index(p.z, a);
  return (iz * a.grid_n + iy) * a.grid_n + ix;
}

// Minimum-image displacement of the walker from object o (geometries.py:2067-2068).
template <int DIM>
__device__ __forceinline__ void min_image(const Pos& p, const float4 o, const WalkArgs& a,
                                          float& qx, float& qy, float& qz) {
  const float L = a.cell;
  qx = p.x - o.x; qx = qx - L * img_shift(qx, a);
  qy = p.y - o.y; qy = qy - L * img_shift(qy, a);
  if (DIM == 3) { qz = p.z - o.z; qz = qz - L * img_shift(qz, a); }
  else qz = 0.0f;
}

template <int DIM>
__device__ __forceinline__ float dot_d(float ax, float ay, float az, float bx, float by, float bz) {
  if (DIM == 2) return ax * bx + ay * by;
  return (ax * bx + ay * by) + az * bz;
}

// Exterior reflection off the nearest object + fold-first clamp pass
// (PackedCylinders.reflect / reflect_with_log_weight geometries.py:1746-2016,
// PackedSpheres geometries.py:2445-2631).  Returns d_perp of the hit (0 if none).
template <int DIM>
__device__ __forceinline__ float packed_reflect(Pos& p, float sx, float sy, float sz, const WalkArgs& a,
                                                const Tables& tb, uint32_t& hits, uint32_t& ncand) {
  const float EPS = a.eps, NUDGE = a.nudge, L = a.cell;
  if (DIM == 2) (void)sz;
  const float sl = (DIM == 2) ? sqrt_rn(sx * sx + sy * sy) : sqrt_rn((sx * sx + sy * sy) + sz * sz);
  float dx = 0.0f, dy = 0.0f, dz = 0.0f;
  if (sl > 0.0f) { dx = div_rn(sx, sl); dy = div_rn(sy, sl); if (DIM == 3) dz = div_rn(sz, sl); }
  const int gc = cell_of<DIM>(p, a);
  const int beg = tb.start[gc], end = tb.start[gc + 1];
  ncand += (uint32_t)(end - beg);
  float best_t = __builtin_inff(), best_disc = 0.0f, bqx = 0.0f, bqy = 0.0f, bqz = 0.0f, bR = 1.0f;
  bool hit = false;
  for (int i = beg; i < end; ++i) {
    const float4 o = tb.objs[tb.items[i]];
    float qx, qy, qz;
    min_image<DIM>(p, o, a, qx, qy, qz);
    const float RR = o.w * o.w;
    const float dp = dot_d<DIM>(dx, dy, dz, qx, qy, qz);
    const float disc = dp * dp - (dot_d<DIM>(qx, qy, qz, qx, qy, qz) - RR);
    const float t = -dp - sqrt_rn(fmaxf(disc, 0.0f));
    const bool valid = (disc > 0.0f) && (t > EPS) && (t < sl) && (sl > 0.0f);
    if (valid && t < best_t) {
      best_t = t; best_disc = disc; bqx = qx; bqy = qy; bqz = qz; bR = o.w; hit = true;
    }
  }
  float fx = p.x + sx, fy = p.y + sy, fz = p.z + (DIM == 3 ? sz : 0.0f);
  float dperp = 0.0f;
  if (hit) {
    hits++;
    const float ts = best_t;
    const float qhx = bqx + ts * dx, qhy = bqy + ts * dy, qhz = bqz + ts * dz;
    const float nx = div_rn(qhx, bR), ny = div_rn(qhy, bR), nz = (DIM == 3) ? div_rn(qhz, bR) : 0.0f;
    const float tdd = 2.0f * dot_d<DIM>(dx, dy, dz, nx, ny, nz);
    float rx = dx - tdd * nx, ry = dy - tdd * ny, rz = (DIM == 3) ? dz - tdd * nz : 0.0f;
    const float rn = sqrt_rn(dot_d<DIM>(rx, ry, rz, rx, ry, rz));
    if (rn > 0.0f) {
      const float rs = fmaxf(rn, 1e-30f);
      rx = div_rn(rx, rs); ry = div_rn(ry, rs); rz = div_rn(rz, rs);
    } else {
      rx = 0.0f; ry = 0.0f; rz = 0.0f;
    }
    const float remaining = sl - ts;
    const float go = fmaxf(remaining - NUDGE, 0.0f);
    fx = ((p.x + ts * dx) + NUDGE * nx) + rx * go;
    fy = ((p.y + ts * dy) + NUDGE * ny) + ry * go;
    if (DIM == 3) fz = ((p.z + ts * dz) + NUDGE * nz) + rz * go;
    dperp = remaining * div_rn(sqrt_rn(fmaxf(best_disc, 0.0f)), bR);
  }
  // Safety clamp on the folded end point, over the same candidates
  // (geometries.py:1858-1877).
  const float gx = fx - L * fold_shift(fx, a);
  const float gy = fy - L * fold_shift(fy, a);
  const float gz = (DIM == 3) ? fz - L * fold_shift(fz, a) : 0.0f;
  float best_pen = 1.0f, cqx = 0.0f, cqy = 0.0f, cqz = 0.0f, cR = 1.0f;
  for (int i = beg; i < end; ++i) {
    const float4 o = tb.objs[tb.items[i]];
    float qx = gx - o.x; qx = qx - L * img_shift(qx, a);
    float qy = gy - o.y; qy = qy - L * img_shift(qy, a);
    float qz = 0.0f;
    if (DIM == 3) { qz = gz - o.z; qz = qz - L * img_shift(qz, a); }
    const float d2 = dot_d<DIM>(qx, qy, qz, qx, qy, qz);
    const float RR = o.w * o.w;
    if (d2 < RR) {
      const float pen = div_rn(d2, RR);
      if (pen < best_pen) { best_pen = pen; cqx = qx; cqy = qy; cqz = qz; cR = o.w; }
    }
  }
  if (best_pen < 1.0f) {
    const float dcl = sqrt_rn(dot_d<DIM>(cqx, cqy, cqz, cqx, cqy, cqz));
    const float scale = div_rn(cR + NUDGE, fmaxf(dcl, NUDGE));
    const float sm1 = scale - 1.0f;
    fx = fx + cqx * sm1;
    fy = fy + cqy * sm1;
    if (DIM == 3) fz = fz + cqz * sm1;
  }
  p.x = fx; p.y = fy;
  p.z = (DIM == 3) ? fz : p.z + sz;
  return dperp;
}

// Powles permeation against the nearest object wall, from inside or outside
// (PackedCylinders.permeate geometries.py:2018-2183, PackedSpheres 2633-2756).
// The candidate objects are the list [beg, end) of a grid cell whose listed
// objects include every object within one step of the walker; with MASKED, only
// the entries beg + j with bit j of `mask` set (the others are known to have no
// valid root), in list order.
template <int DIM, bool MASKED = false>
__device__ __forceinline__ float packed_permeate_list(Pos& p, float sx, float sy, float sz, const WalkArgs& a,
                                                      const Tables& tb, int beg, int end, const PowlesSrc& ps,
                                                      uint32_t& hits, uint32_t& ncross, uint32_t& ncand,
                                                      uint32_t mask = 0u) {
  const float EPS = a.eps, NUDGE = a.nudge;
  const float sl = (DIM == 2) ? sqrt_rn(sx * sx + sy * sy) : sqrt_rn((sx * sx + sy * sy) + sz * sz);
  float dx = 0.0f, dy = 0.0f, dz = 0.0f;
  if (sl > 0.0f) { dx = div_rn(sx, sl); dy = div_rn(sy, sl); if (DIM == 3) dz = div_rn(sz, sl); }
  ncand += MASKED ? (uint32_t)__builtin_popcount(mask) : (uint32_t)(end - beg);
  float best_t = __builtin_inff(), best_disc = 0.0f, bqx = 0.0f, bqy = 0.0f, bqz = 0.0f, bR = 1.0f;
  bool hit = false, hit_inside = false;
  for (int i = MASKED ? beg + __builtin_ctz(mask | 0x80000000u) : beg; MASKED ? mask != 0u : i < end;
       i = MASKED ? beg + __builtin_ctz(mask | 0x80000000u) : i + 1) {
    if (MASKED) mask &= mask - 1u;
    const float4 o = tb.objs[tb.items[i]];
    float qx, qy, qz;
    min_image<DIM>(p, o, a, qx, qy, qz);
    const float RR = o.w * o.w;
    const float dist2 = dot_d<DIM>(qx, qy, qz, qx, qy, qz);
    const bool inside = dist2 < RR;
    const float dp = dot_d<DIM>(dx, dy, dz, qx, qy, qz);
    const float disc = dp * dp - (dist2 - RR);
    const float sq = sqrt_rn(fmaxf(disc, 0.0f));
    const float t = inside ? (-dp + sq) : (-dp - sq);
    const bool valid = (disc > 0.0f) && (t > EPS) && (t < sl) && (sl > 0.0f);
    if (valid && t < best_t) {
      best_t = t; best_disc = disc; bqx = qx; bqy = qy; bqz = qz; bR = o.w; hit = true;
      hit_inside = inside;
    }
  }
  if (!hit) {
    p.x = p.x + sx; p.y = p.y + sy; p.z = p.z + sz;
    return 0.0f;
  }
  hits++;
  const float ts = best_t;
  const float hx = bqx + ts * dx, hy = bqy + ts * dy, hz = (DIM == 3) ? bqz + ts * dz : 0.0f;
  const float hl = fmaxf(sqrt_rn(dot_d<DIM>(hx, hy, hz, hx, hy, hz)), 1e-30f);
  const float nx = div_rn(hx, hl), ny = div_rn(hy, hl), nz = (DIM == 3) ? div_rn(hz, hl) : 0.0f;
  const float remaining = sl - ts;
  const float dperp = remaining * div_rn(sqrt_rn(fmaxf(best_disc, 0.0f)), bR);
  const float pr = fminf(1.0f, (2.0f * a.kappaD) * dperp);
  if (ps.get() < pr) {
    p.x = p.x + sx; p.y = p.y + sy; p.z = p.z + sz; ncross++;
    return 0.0f;
  }
  const float tdd = 2.0f * dot_d<DIM>(dx, dy, dz, nx, ny, nz);
  float rx = dx - tdd * nx, ry = dy - tdd * ny, rz = (DIM == 3) ? dz - tdd * nz : 0.0f;
  const float rn = sqrt_rn(dot_d<DIM>(rx, ry, rz, rx, ry, rz));
  if (rn > 0.0f) {
    const float rs = fmaxf(rn, 1e-30f);
    rx = div_rn(rx, rs); ry = div_rn(ry, rs); rz = div_rn(rz, rs);
  } else {
    rx = 0.0f; ry = 0.0f; rz = 0.0f;
  }
  const float shx = bR * nx, shy = bR * ny, shz = bR * nz;
  const float ux = hit_inside ? shx - NUDGE * nx : shx + NUDGE * nx;
  const float uy = hit_inside ? shy - NUDGE * ny : shy + NUDGE * ny;
  const float uz = hit_inside ? shz - NUDGE * nz : shz + NUDGE * nz;
  const float go = fmaxf(remaining - NUDGE, 0.0f);
  float fx = ux + rx * go, fy = uy + ry * go, fz = (DIM == 3) ? uz + rz * go : 0.0f;
  const float fn = sqrt_rn(dot_d<DIM>(fx, fy, fz, fx, fy, fz));
  const float fs = fmaxf(fn, 1e-30f);
  const float target = hit_inside ? bR - NUDGE : bR + NUDGE;
  const bool wrong = hit_inside ? (fn >= bR) : (fn <= bR);
  if (wrong) { fx = div_rn(fx * target, fs); fy = div_rn(fy * target, fs); if (DIM == 3) fz = div_rn(fz * target, fs); }
  p.x = p.x + (fx - bqx);
  p.y = p.y + (fy - bqy);
  p.z = (DIM == 3) ? p.z + (fz - bqz) : p.z + sz;
  return (-2.0f * a.rhoD) * dperp;
}

template <int DIM>
__device__ __forceinline__ float packed_permeate(Pos& p, float sx, float sy, float sz, const WalkArgs& a,
                                                 const Tables& tb, const PowlesSrc& ps, uint32_t& hits,
                                                 uint32_t& ncross, uint32_t& ncand) {
  const int gc = cell_of<DIM>(p, a);
  return packed_permeate_list<DIM>(p, sx, sy, sz, a, tb, tb.start[gc], tb.start[gc + 1], ps, hits, ncross, ncand);
}

// ---------------------------------------------------------------------------
// Screened packed permeation.  Each lane keeps the grid cell it walks in, as a
// box expanded on every side by a.scr_w, or less for the few cells whose list
// would not fit the cache (the host lists, for that box, every object within one
// step of it), and, when the list holds at most SCR_MAXC objects, each object's
// image centre and the side of its wall the walker is clear on (screen.h: the
// pair rs, b).  A sub-step that ends inside the box and passes the screen for
// every cached object has no valid root in the reference's test
// (geometries.py:2067-2093, 2667-2691): its result is r + s, computed here
// exactly as the reference's no-hit branch does (geometries.py:2099-2102), and
// the screen state stays as it is.  Only the other lanes run the reference's
// root test, over the cached cell's list, and then set their state anew.
// ---------------------------------------------------------------------------
template <int DIM>
struct PackedScreen {
  int beg, len;            // the cell's list in LDS
  float blo[DIM], bhi[DIM];
  // per cached object: image centre, rs = -R (clear inside) or +R, and the pass
  // term b (screen.h); unused slots always pass (b = +inf), and slot 0 of a lane
  // whose list is not cached never passes (b = -inf)
  float cx[SCR_MAXC], cy[SCR_MAXC], cz[SCR_MAXC], rs[SCR_MAXC], b[SCR_MAXC];

  __device__ __forceinline__ bool uncached(const WalkArgs& a) const { return !a.scr_cache || len > SCR_MAXC; }

  __device__ __forceinline__ bool in_box(float x, float y, float z) const {
    bool in = (x > blo[0]) & (x < bhi[0]) & (y > blo[1]) & (y < bhi[1]);
    if (DIM == 3) in = in & (z > blo[DIM - 1]) & (z < bhi[DIM - 1]);
    return in;
  }

  // Slot j's state at p for an object of radius R (unused and uncached slots as above).
  __device__ __forceinline__ void slot_state(int j, const Pos& p, float R, const WalkArgs& a) {
    const int nc = uncached(a) ? 0 : len;
    float r, bb;
    screen_state(screen_d2(p.x, p.y, p.z, cx[j], cy[j], cz[j], DIM), R, a.scr_m2, a.scr_cout, a.scr_cin, r, bb);
    rs[j] = j < nc ? r : 1.0f;
    b[j] = j < nc ? bb : ((j == 0 && uncached(a)) ? -__builtin_inff() : __builtin_inff());
  }

  // The state of every slot at p (same cell, images and radii) after the
  // reference's test moved, reflected or stopped the walker.  A slot whose state
  // the new position still passes keeps it (the position is clear on the same side
  // of that wall); the others are set anew, skipped by the wavefront when none of
  // its lanes needs it.
  __device__ __forceinline__ void set_state(const Pos& p, const WalkArgs& a) {
    static_for<SCR_MAXC>([&](auto jc) {
      constexpr int j = decltype(jc)::value;
      const bool keep = screen_pass(screen_d2(p.x, p.y, p.z, cx[j], cy[j], cz[j], DIM), rs[j], b[j]);
      if (__ballot(!keep) != 0ull && !keep) slot_state(j, p, fabsf(rs[j]), a);
    });
  }

  // The walker's cell (cell_of), the list and box level the cell's list start
  // carries, its box in the walker's frame (the cell expanded by that level's
  // width), the cached objects' nearest images and the walker's state.
  __device__ __forceinline__ void refresh(const Pos& p, const WalkArgs& a, const Tables& tb) {
    const float L = a.cell, h = a.grid_h;
    const int gc = cell_of<DIM>(p, a);
    const unsigned s0 = (unsigned)tb.start[gc];
    beg = (int)(s0 & SCR_LIST_MASK);
    len = (int)((unsigned)tb.start[gc + 1] & SCR_LIST_MASK) - beg;
    const float w = screen_box_width(a.scr_w, s0 >> SCR_LEVEL_SHIFT);
    const float v[3] = {p.x, p.y, p.z};
    static_for<DIM>([&](auto dc) {
      constexpr int d = decltype(dc)::value;
      const float k = floorf(v[d] * a.inv_cell + 0.5f);
      const int i = grid_index(v[d], a);
      const float lo = (-0.5f * L + (float)i * h) + L * k;
      blo[d] = lo - w;
      bhi[d] = (lo + h) + w;
    });
    const int nc = uncached(a) ? 0 : len;
    // every slot is written (unused ones with zeros), so no earlier value stays live
    static_for<SCR_MAXC>([&](auto jc) {
      constexpr int j = decltype(jc)::value;
      const bool use = j < nc;
      const float4 o = use ? tb.objs[tb.items[beg + j]] : float4{0.0f, 0.0f, 0.0f, 0.0f};
      cx[j] = use ? o.x + L * floorf((v[0] - o.x) * a.inv_cell + 0.5f) : 0.0f;
      cy[j] = use ? o.y + L * floorf((v[1] - o.y) * a.inv_cell + 0.5f) : 0.0f;
      cz[j] = (DIM == 3 && use) ? o.z + L * floorf((v[2] - o.z) * a.inv_cell + 0.5f) : 0.0f;
      slot_state(j, p, o.w, a);
    });
  }

  // True when the sub-step to (nx, ny, nz) = p + s passes for every cached object;
  // `leave` is set when the end point is outside the cell box.  Branch-free.
  __device__ __forceinline__ bool pass(float nx, float ny, float nz, bool& leave) const {
    bool ok = true;
    static_for<SCR_MAXC>([&](auto jc) {
      constexpr int j = decltype(jc)::value;
      ok = ok & screen_pass(screen_d2(nx, ny, nz, cx[j], cy[j], cz[j], DIM), rs[j], b[j]);
    });
    leave = !in_box(nx, ny, nz);
    return ok;
  }

  // The cached objects (bit j for slot j) the sub-step to (nx, ny, nz) does not pass.
  __device__ __forceinline__ uint32_t flags(float nx, float ny, float nz) const {
    uint32_t f = 0u;
    static_for<SCR_MAXC>([&](auto jc) {
      constexpr int j = decltype(jc)::value;
      f |= (uint32_t)!screen_pass(screen_d2(nx, ny, nz, cx[j], cy[j], cz[j], DIM), rs[j], b[j]) << j;
    });
    return f;
  }
};

// One screened sub-step of a permeable packing (the wavefront's lanes converged):
// r + s, the reference's no-hit result, unless the screen cannot pass it; then the
// reference's root test over the objects the screen did not pass and the walker's
// new screen state.  Returns the sub-step's unit log-weight.
template <int DIM>
__device__ __forceinline__ float screened_permeate(Pos& p, float sx, float sy, float sz, const WalkArgs& a,
                                                   const Tables& tb, PackedScreen<DIM>& scr,
                                                   const PowlesSrc& ps, uint32_t& hits, uint32_t& ncross,
                                                   uint32_t& ncand) {
  const float nx = p.x + sx, ny = p.y + sy, nz = p.z + sz;
  bool leave;
  const bool exact = !scr.pass(nx, ny, nz, leave);
  float dlog = 0.0f;
  if (__ballot(exact | leave) == 0ull) {
    p.x = nx; p.y = ny; p.z = nz;
  } else if (!a.scr_cache) {
    // no cached objects (lists too long to mask): the reference's test, every lane
    dlog = packed_permeate_list<DIM>(p, sx, sy, sz, a, tb, scr.beg, scr.beg + scr.len, ps, hits, ncross,
                                     ncand);
    if (!scr.in_box(p.x, p.y, p.z)) scr.refresh(p, a, tb);
  } else {
    if (exact) {
      // the reference's root test over the objects the screen did not pass
      // (the whole list when it is not cached)
      const uint32_t m = scr.uncached(a) ? (uint32_t)((1ull << scr.len) - 1ull) : scr.flags(nx, ny, nz);
      dlog = packed_permeate_list<DIM, true>(p, sx, sy, sz, a, tb, scr.beg, 0, ps, hits, ncross, ncand, m);
    } else {
      p.x = nx; p.y = ny; p.z = nz;
    }
    if (!scr.in_box(p.x, p.y, p.z)) scr.refresh(p, a, tb);   // a new cell
    else if (exact) scr.set_state(p, a);                     // moved, reflected or stopped
  }
  return dlog;
}

// ---------------------------------------------------------------------------
// PackedMyelinatedCylinders (make_packed_myelin_traj_step_fn, physics.py:682-978,
// without magnetisation transfer, as the fused forward calls it).  Walker
// compartment: Pos::cx = class (0 intra, 1 myelin, 2 extra), Pos::cy = axon
// (0 for extra-axonal walkers, whose membrane tests use axon 0 as the reference's
// k_cyl does), Pos::cz = start class.  LDS rows per axon k: objs[k] = (cx, cy,
// R_in, R_out), objs[n + k] = step lengths (intra, myelin, extra), objs[2n + k] =
// 2 kappa/D (inner from intra, inner from myelin, outer from myelin, outer from
// extra).  Minimum images use true division, as the reference does.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float pm_wrap(float q, float L) { return q - L * floorf(div_rn(q, L) + 0.5f); }

__device__ __forceinline__ Pos pm_point(float x, float y) { return Pos{x, y, 0.0f, 0, 0, 0}; }

// The exterior clamp's target: argmin over the axons listed for the cell of (x, y)
// of d^2/R_out^2 among axons whose outer disc contains the point (first index on
// ties; -1 when none does), physics.py:872-882.
__device__ __forceinline__ int pm_nearest_container(float x, float y, const WalkArgs& a,
                                                    const Tables& tb, uint32_t& ncand) {
  const float L = a.cell;
  const int gc = cell_of<2>(pm_point(x, y), a);
  const int beg = tb.start[gc], end = tb.start[gc + 1];
  ncand += (uint32_t)(end - beg);
  float best = 1.0f;
  int kb = -1;
  for (int i = beg; i < end; ++i) {
    const int j = tb.items[i];
    const float4 g = tb.objs[j];
    const float fx = pm_wrap(x - g.x, L), fy = pm_wrap(y - g.y, L);
    const float d2 = fx * fx + fy * fy;
    const float RR = g.w * g.w;
    const float pen = (g.w > 0.0f) ? (d2 < RR ? div_rn(d2, RR + 1e-30f) : 1.0f) : 1.0f;
    if (pen < best) { best = pen; kb = j; }
  }
  return kb;
}

// One sub-step; returns the unit (rho/D = 1) boundary log-weight.
__device__ __forceinline__ float pmyel_step(Pos& p, float ux, float uy, float uz, const WalkArgs& a,
                                            const Tables& tb, const PowlesSrc& ps, uint32_t& hits,
                                            uint32_t& ncross, uint32_t& ncand) {
  const int n = a.n_obj;
  const float L = a.cell, NUDGE = a.nudge;
  const int cls = p.cx, k = p.cy;
  const float4 geo = tb.objs[k];
  const float4 stp = tb.objs[n + k];
  const float4 k2 = tb.objs[2 * n + k];
  const bool is_intra = cls == 0, is_myelin = cls == 1, is_extra = cls == 2;
  const float sl = is_intra ? stp.x : (is_myelin ? stp.y : stp.z);
  const float sx = ux * sl, sy = uy * sl, sz = uz * sl;
  const float Rin = geo.z, Rout = geo.w;
  float qx = pm_wrap((p.x + sx) - geo.x, L), qy = pm_wrap((p.y + sy) - geo.y, L);
  const float nrm = sqrt_rn(qx * qx + qy * qy);
  const bool cio = is_intra && nrm >= Rin, cii = is_myelin && nrm < Rin;
  const bool coo = is_myelin && nrm >= Rout, coi = is_extra && nrm < Rout;
  const bool ci = cio || cii, co = coo || coi;
  const float dpi = fabsf(nrm - Rin), dpo = fabsf(nrm - Rout);
  bool t_in = false, t_out = false;
  if (ci || co) {
    hits++;
    const float u = ps.get();
    if (ci) t_in = u < fminf(1.0f, (is_intra ? k2.x : k2.y) * dpi);
    else t_out = u < fminf(1.0f, (is_myelin ? k2.z : k2.w) * dpo);
  }
  int ncls = cls, nk = k;
  float dlog = 0.0f;
  // inner membrane: transmit, or mirror radially to 2 R_in - |q| (physics.py:812-830)
  if (ci) {
    if (t_in) {
      ncls = cio ? 1 : 0;
    } else {
      const float s = fmaxf(nrm, 1e-20f);
      const float hx = div_rn(qx, s), hy = div_rn(qy, s);
      const float rr = 2.0f * Rin - nrm;
      qx = hx * rr;
      qy = hy * rr;
      dlog = dlog + (-2.0f * dpi);
    }
  }
  // outer membrane, on the recomputed norm (physics.py:832-851)
  if (co) {
    if (t_out) {
      ncls = coo ? 2 : 1;
      nk = coo ? 0 : k;
    } else {
      const float n2 = sqrt_rn(qx * qx + qy * qy);
      const float s = fmaxf(n2, 1e-20f);
      const float hx = div_rn(qx, s), hy = div_rn(qy, s);
      const float rr = 2.0f * Rout - n2;
      qx = hx * rr;
      qy = hy * rr;
      dlog = dlog + (-2.0f * dpo);
    }
  }
  if (t_in || t_out) ncross++;
  float xa = pm_wrap(qx + geo.x, L), ya = pm_wrap(qy + geo.y, L);
  // extra-axonal walkers: mirror off the outer wall of the nearest containing axon
  // (physics.py:853-890)
  if (is_extra) {
    const int kc = pm_nearest_container(xa, ya, a, tb, ncand);
    if (kc >= 0) {
      const float4 gc = tb.objs[kc];
      const float cqx = pm_wrap(xa - gc.x, L), cqy = pm_wrap(ya - gc.y, L);
      const float dcl = sqrt_rn(cqx * cqx + cqy * cqy);
      const float cnx = xa - cqx, cny = ya - cqy;
      const float drefl = 2.0f * gc.w - dcl;
      const float den = fmaxf(dcl, NUDGE);
      xa = cnx + div_rn(cqx * drefl, den);
      ya = cny + div_rn(cqy * drefl, den);
      dlog = dlog + (-2.0f * fmaxf(gc.w - dcl, 0.0f));
    }
  }
  // wrong-side clamps of untransmitted intra / myelin walkers (physics.py:892-916)
  const bool same = !(t_in || t_out);
  if (same && (is_intra || is_myelin)) {
    float ex = pm_wrap(xa - geo.x, L), ey = pm_wrap(ya - geo.y, L);
    float de = sqrt_rn(ex * ex + ey * ey);
    if (is_intra && de >= Rin) {
      const float s = fmaxf(de, 1e-20f);
      xa = geo.x + (div_rn(ex, s)) * (Rin - NUDGE);
      ya = geo.y + (div_rn(ey, s)) * (Rin - NUDGE);
    }
    if (is_myelin && de < Rin) {
      const float s = fmaxf(de, 1e-20f);
      xa = geo.x + (div_rn(ex, s)) * (Rin + NUDGE);
      ya = geo.y + (div_rn(ey, s)) * (Rin + NUDGE);
      ex = pm_wrap(xa - geo.x, L);
      ey = pm_wrap(ya - geo.y, L);
      de = sqrt_rn(ex * ex + ey * ey);
    }
    if (is_myelin && de >= Rout) {
      const float s = fmaxf(de, 1e-20f);
      xa = geo.x + (div_rn(ex, s)) * (Rout - NUDGE);
      ya = geo.y + (div_rn(ey, s)) * (Rout - NUDGE);
    }
  }
  // intra / myelin walkers are re-classified against their own axon (physics.py:926-949)
  if (!is_extra) {
    const float fx = pm_wrap(xa - geo.x, L), fy = pm_wrap(ya - geo.y, L);
    const float d2 = fx * fx + fy * fy;
    if (d2 < Rin * Rin) { ncls = 0; nk = k; }
    else if (d2 < Rout * Rout) { ncls = 1; nk = k; }
  }
  p.x = xa;
  p.y = ya;
  p.z = p.z + sz;
  p.cx = ncls;
  p.cy = nk;
  return dlog;
}

// Compartment of a position (classify_position): 0/1 for Cylinder and Sphere
// (geometries.py:180-184, 803-813), 0 or the largest 1-based index of a
// containing object for packed substrates (geometries.py:2185-2214, 2758-2771).
template <int GEOM>
__device__ __forceinline__ int classify(const Pos& p, const WalkArgs& a, const Tables& tb) {
  if (GEOM == G_CYL) {
    const float RR = a.size * a.size;
    return (p.x * p.x + p.y * p.y) < RR ? 0 : 1;
  } else if (GEOM == G_SPH) {
    const float RR = a.size * a.size;
    return ((p.x * p.x + p.y * p.y) + p.z * p.z) < RR ? 0 : 1;
  } else if (GEOM == G_ELL) {
    return ell_form(p.x, p.y, p.z, p.x, p.y, p.z, a) < 1.0f ? 0 : 1;
  } else if (GEOM == G_SLAB) {
    return p.x < 0.5f * a.size ? 0 : 1;
  } else if (GEOM == G_SHELL || GEOM == G_SHELLC) {
    return shell_norm<GEOM == G_SHELLC>(p.x, p.y, p.z) < a.size ? 0 : 1;
  } else if (GEOM == G_MYEL || GEOM == G_PMYEL) {
    return p.cx;  // the carried compartment (core.py:486-530; packed: class, core.py:549-570)
  } else if (GEOM == G_PCYL || GEOM == G_PSPH) {
    constexpr int DIM = (GEOM == G_PCYL) ? 2 : 3;
    const int gc = cell_of<DIM>(p, a);
    int id = 0;
    // (the screened walk's list starts carry the cell's box level above the mask)
    const int end = (int)((unsigned)tb.start[gc + 1] & SCR_LIST_MASK);
    for (int i = (int)((unsigned)tb.start[gc] & SCR_LIST_MASK); i < end; ++i) {
      const int k = tb.items[i];
      const float4 o = tb.objs[k];
      float qx, qy, qz;
      min_image<DIM>(p, o, a, qx, qy, qz);
      if (dot_d<DIM>(qx, qy, qz, qx, qy, qz) < o.w * o.w) id = k + 1 > id ? k + 1 : id;
    }
    return id;
  }
  return 0;
}

// Fold a packed position into the primary cell, carrying the cell index.
template <int DIM>
__device__ __forceinline__ void fold(Pos& p, const WalkArgs& a) {
  const float L = a.cell;
  float k = fold_shift(p.x, a);
  if (k != 0.0f) { p.x = p.x - L * k; p.cx += (int)k; }
  k = fold_shift(p.y, a);
  if (k != 0.0f) { p.y = p.y - L * k; p.cy += (int)k; }
  if (DIM == 3) {
    k = fold_shift(p.z, a);
    if (k != 0.0f) { p.z = p.z - L * k; p.cz += (int)k; }
  }
}

// One sub-step of wall physics for any geometry; returns the log-weight
// increment before the chi gating (physics.py:303-311).
template <int GEOM, int WALL>
__device__ __forceinline__ float wall_step(Pos& p, float sx, float sy, float sz, const WalkArgs& a,
                                           const Tables& tb, const PowlesSrc& ps, uint32_t& hits,
                                           uint32_t& ncross, uint32_t& ncand) {
  constexpr int DIM = (GEOM == G_PSPH) ? 3 : 2;
  float dlog = 0.0f;
  if (GEOM == G_FREE) {
    p.x = p.x + sx; p.y = p.y + sy; p.z = p.z + sz;
  } else if (GEOM == G_BOX1D) {
    dlog = box_step<WALL>(p, sx, sy, sz, a, hits);
  } else if (GEOM == G_CYL) {
    if (WALL == W_PERMEATE) {
      dlog = cyl_permeate(p.x, p.y, sx, sy, a, ps, hits, ncross);
    } else {
      const float ds = cyl_reflect(p.x, p.y, sx, sy, a, hits);
      if (WALL == W_SURFACE) dlog = (-2.0f * a.rhoD) * ds;
    }
    p.z = p.z + sz;
  } else if (GEOM == G_SPH) {
    if (WALL == W_PERMEATE) {
      dlog = sph_permeate(p, sx, sy, sz, a, ps, hits, ncross);
    } else {
      const float ds = sph_reflect(p, sx, sy, sz, a, hits);
      if (WALL == W_SURFACE) dlog = (-2.0f * a.rhoD) * ds;
    }
  } else if (GEOM == G_ELL) {
    if (WALL == W_PERMEATE) {
      dlog = ell_permeate(p, sx, sy, sz, a, ps, hits, ncross);
    } else {
      const float ds = ell_reflect(p, sx, sy, sz, a, hits);
      if (WALL == W_SURFACE) dlog = (-2.0f * a.rhoD) * ds;
    }
  } else if (GEOM == G_SLAB) {
    dlog = slab_step<WALL>(p, sx, sy, sz, a, ps, hits, ncross);
  } else if (GEOM == G_SHELL || GEOM == G_SHELLC) {
    dlog = shell_step<GEOM == G_SHELLC, WALL>(p, sx, sy, sz, a, ps, hits, ncross);
  } else if (GEOM == G_MYEL) {
    myel_step(p, sx, sy, sz, a, ps, hits, ncross);
  } else {
    if (WALL == W_PERMEATE) {
      dlog = packed_permeate<DIM>(p, sx, sy, sz, a, tb, ps, hits, ncross, ncand);
    } else {
      const float dp = packed_reflect<DIM>(p, sx, sy, sz, a, tb, hits, ncand);
      if (WALL == W_SURFACE) dlog = (-2.0f * a.rhoD) * dp;
    }
  }
  return dlog;
}

// ---------------------------------------------------------------------------
// Start positions sampled in-kernel (reference: host rejection sampling,
// geometries.py:87-91, 186-195, 477-491, 1717-1744, 2421-2443).
// ---------------------------------------------------------------------------
template <int GEOM>
__device__ __forceinline__ Pos sample_start(uint32_t w, const WalkArgs& a, const Tables& tb) {
  Pos p{0.0f, 0.0f, 0.0f, 0, 0, 0};
  if (GEOM == G_FREE) return p;
  // MyelinatedCylinder: the start compartment is fixed by the walker index (the
  // reference allots round(N V_c / V) walkers per compartment in index order,
  // extra-axonal space bounded at 2 R_out, geometries.py:921-1008).
  const int c0 = (GEOM == G_MYEL) ? ((long long)w < a.m_n0 ? 0 : ((long long)w < a.m_n01 ? 1 : 2)) : 0;
  const float r_hi = (GEOM == G_MYEL) ? pick3(c0, a.size, a.aux, 2.0f * a.aux) : 0.0f;
  p.cx = c0;
  p.cy = c0;
  // PackedMyelinatedCylinders: segment j of the start boundaries is intra of axon j
  // (j < n), myelin of axon j - n (j < 2n), extra-axonal past the last boundary
  // (geometries.py:3112-3210: area allotment in this order).
  int pm_cls = 2, pm_k = 0;
  if (GEOM == G_PMYEL) {
    const int n2 = 2 * a.n_obj;
    if ((long long)w < a.pm_start[n2]) {
      int lo = 0, hi = n2;                       // first j with start[j] > w, minus 1
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (a.pm_start[mid] <= (long long)w) lo = mid; else hi = mid;
      }
      pm_cls = lo < a.n_obj ? 0 : 1;
      pm_k = lo < a.n_obj ? lo : lo - a.n_obj;
    }
    p.cx = pm_cls;
    p.cy = pm_k;
    p.cz = pm_cls;
  }
  for (uint32_t attempt = 0; attempt < 4096u; ++attempt) {
    const u32x4 r = philox4x32_10({w, a.instance, attempt, TAG_INIT}, a.seed_lo, a.seed_hi);
    const float u0 = u01(r.x), u1 = u01(r.y), u2 = u01(r.z);
    if (GEOM == G_BOX1D) {
      p.x = u0 * a.size;
      return p;
    } else if (GEOM == G_CYL) {
      p.x = (2.0f * u0 - 1.0f) * a.size;
      p.y = (2.0f * u1 - 1.0f) * a.size;
      if (p.x * p.x + p.y * p.y < a.size * a.size) return p;
    } else if (GEOM == G_SPH || GEOM == G_SHELL) {
      p.x = (2.0f * u0 - 1.0f) * a.size;
      p.y = (2.0f * u1 - 1.0f) * a.size;
      p.z = (2.0f * u2 - 1.0f) * a.size;
      if ((p.x * p.x + p.y * p.y) + p.z * p.z < a.size * a.size) return p;
    } else if (GEOM == G_SHELLC) {
      p.x = (2.0f * u0 - 1.0f) * a.size;
      p.y = (2.0f * u1 - 1.0f) * a.size;
      p.z = u2 * a.size;
      if (p.x * p.x + p.y * p.y < a.size * a.size) return p;
    } else if (GEOM == G_SLAB) {
      p.x = u0 * (0.5f * a.size);
      return p;
    } else if (GEOM == G_MYEL) {
      p.x = (2.0f * u0 - 1.0f) * r_hi;
      p.y = (2.0f * u1 - 1.0f) * r_hi;
      if (myel_classify(p.x, p.y, a) == c0 && p.x * p.x + p.y * p.y < r_hi * r_hi) return p;
    } else if (GEOM == G_PMYEL) {
      if (pm_cls == 2) {
        // uniform in the cell, outside every outer disc
        p.x = (u0 - 0.5f) * a.cell;
        p.y = (u1 - 0.5f) * a.cell;
        bool outside = true;
        const int gc = cell_of<2>(p, a);
        for (int i = tb.start[gc]; i < tb.start[gc + 1]; ++i) {
          const float4 g = tb.objs[tb.items[i]];
          const float fx = pm_wrap(p.x - g.x, a.cell), fy = pm_wrap(p.y - g.y, a.cell);
          outside = outside && (fx * fx + fy * fy > g.w * g.w);
        }
        if (outside) return p;
      } else {
        const float4 g = tb.objs[pm_k];
        const float rh = pm_cls == 0 ? g.z : g.w;
        const float qx = (2.0f * u0 - 1.0f) * rh, qy = (2.0f * u1 - 1.0f) * rh;
        const float d2 = qx * qx + qy * qy;
        const bool ok = pm_cls == 0 ? d2 < rh * rh : (d2 >= g.z * g.z && d2 < rh * rh);
        if (ok) {
          p.x = qx + g.x;
          p.y = qy + g.y;
          return p;
        }
      }
    } else if (GEOM == G_ELL) {
      // unit ball scaled by the semi-axes (uniform: constant Jacobian)
      const float bx = 2.0f * u0 - 1.0f, by = 2.0f * u1 - 1.0f, bz = 2.0f * u2 - 1.0f;
      p.x = bx / sqrtf(a.ax);
      p.y = by / sqrtf(a.ay);
      p.z = bz / sqrtf(a.az);
      if ((bx * bx + by * by) + bz * bz < 1.0f && classify<GEOM>(p, a, tb) == 0) return p;
    } else {
      p.x = (u0 - 0.5f) * a.cell;
      p.y = (u1 - 0.5f) * a.cell;
      p.z = (GEOM == G_PSPH) ? (u2 - 0.5f) * a.cell : 0.0f;
      if (classify<GEOM>(p, a, tb) == 0) return p;
    }
  }
  return p;
}

// float32 -> float16 bits, round to nearest even (numpy's astype(float16)).
__device__ __forceinline__ uint16_t f16_bits(float x) { return __builtin_bit_cast(uint16_t, (_Float16)x); }

// Compartment id of the walk-once producer for permeable walls
// (dmipy_sim/core.py:1166-1185, 1231-1245): the rule keys on the geometry's
// `radius` attribute, so a Cylinder is outside iff its axial-plane norm reaches
// R, and every other geometry with a radius (Sphere; PermeableSlab1D, radius
// L/2; PermeableShell of either kind, radius R_in) iff the 3-D norm reaches it.
// An Ellipsoid has no radius: always 0.
template <int GEOM>
__device__ __forceinline__ float outside_norm(const Pos& p, const WalkArgs& a) {
  if (GEOM == G_CYL) return sqrtf(p.x * p.x + p.y * p.y) < a.size ? 0.0f : 1.0f;
  const float r3 = sqrtf((p.x * p.x + p.y * p.y) + p.z * p.z);
  if (GEOM == G_SPH || GEOM == G_SHELL || GEOM == G_SHELLC) return r3 < a.size ? 0.0f : 1.0f;
  if (GEOM == G_SLAB) return r3 < 0.5f * a.size ? 0.0f : 1.0f;
  return 0.0f;
}

// sin and cos of a phase: quarter-turn reduction r = x - q*pi/2 with pi/2 split in
// two float32 parts (Cody-Waite, fused multiply-adds), then the Taylor polynomials
// of sin (to r^9) and cos (to r^10) on |r| <= pi/4.  Absolute error ~1e-7 for
// |x| < 1e5 rad.
__device__ __forceinline__ void phase_sincos(float x, float* s, float* c) {
  const float q = rintf(x * 0.636619772f);
  float r = fmaf(-q, 1.57079637f, x);
  r = fmaf(-q, -4.37113883e-8f, r);
  const float r2 = r * r;
  float ps = fmaf(r2, 2.75573192e-6f, -1.98412698e-4f);
  ps = fmaf(r2, ps, 8.33333333e-3f);
  ps = fmaf(r2, ps, -1.66666667e-1f);
  const float sr = fmaf(r * r2, ps, r);
  float pc = fmaf(r2, -2.75573192e-7f, 2.48015873e-5f);
  pc = fmaf(r2, pc, -1.38888889e-3f);
  pc = fmaf(r2, pc, 4.16666667e-2f);
  pc = fmaf(r2, pc, -0.5f);
  const float cr = fmaf(r2, pc, 1.0f);
  // quadrant qi = q mod 4: (cos, sin) = (cr, sr), (-sr, cr), (-cr, -sr), (sr, -cr) --
  // odd quadrants swap the pair, cos flips sign for qi = 1, 2 and sin for qi = 2, 3
  // (exact sign flips, no branches)
  const uint32_t qi = (uint32_t)(int)q;
  const bool swap = (qi & 1u) != 0u;
  const float c0 = swap ? sr : cr, s0 = swap ? cr : sr;
  *c = __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, c0) ^ (((qi + 1u) & 2u) << 30));
  *s = __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, s0) ^ ((qi & 2u) << 30));
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Wave-reduced (Re, Im) of w e^{i(phi - c)} for one measurement block, written
// by lane 0 to dst[m*2 + {0,1}].
template <int MB, bool KAHAN>
__device__ __forceinline__ void emit_partials(const float (&phi)[MB], const float (&kc)[MB], float wgt,
                                              bool active, int lane, double* dst) {
  static_for<MB>([&](auto mc) {
    constexpr int m = decltype(mc)::value;
    float sn, cs;
    phase_sincos(phi[m], &sn, &cs);
    if (KAHAN) {
      const float c0 = cs;
      cs = fmaf(kc[m], sn, cs);
      sn = fmaf(-kc[m], c0, sn);
    }
    const double re = wave_sum(active ? (double)(wgt * cs) : 0.0);
    const double im = wave_sum(active ? (double)(wgt * sn) : 0.0);
    if (lane == 0) { dst[2 * m] = re; dst[2 * m + 1] = im; }
  });
}

// ---------------------------------------------------------------------------
// Phase accumulators.  Per waveform sample the walk adds, for every measurement
// m, dphi = Gs[t][m] . P with P the sum of the n_sub sub-step positions.
// ---------------------------------------------------------------------------

// VALU form: lane = walker, phi[m] (+ Kahan compensation) in VGPRs; G(t) read
// through the constant address space (scalar loads).  Used for small
// measurement blocks and for the compensated form.
template <int MB, bool KAHAN>
struct PhaseVALU {
  float phi[MB], kc[MB];
  __device__ __forceinline__ void init() {
    static_for<MB>([&](auto mc) {
      constexpr int m = decltype(mc)::value;
      phi[m] = 0.0f;
      kc[m] = 0.0f;
    });
  }
  __device__ __forceinline__ void prefetch(const WalkArgs&, int, int, int) {}
  __device__ __forceinline__ void update(const WalkArgs& a, int t, int m0, int, float Px, float Py, float Pz) {
    const cfloat* g = (const cfloat*)(a.Gs + ((size_t)t * a.n_meas_pad + m0) * 3);
    static_for<MB>([&](auto mc) {
      constexpr int m = decltype(mc)::value;
      const float dphi = fmaf(g[3 * m + 2], Pz, fmaf(g[3 * m + 1], Py, g[3 * m] * Px));
      if (KAHAN) {
        const float yk = dphi - kc[m];
        const float tk = phi[m] + yk;
        kc[m] = (tk - phi[m]) - yk;
        phi[m] = tk;
      } else {
        phi[m] = phi[m] + dphi;
      }
    });
  }
  __device__ __forceinline__ void set_ref(const WalkArgs&, int, int, float, float, float) {}
  __device__ __forceinline__ void add_segment(const WalkArgs&, int, int, int) {}
  __device__ __forceinline__ void emit(const WalkArgs&, int, float wgt, bool active, int lane, double* dst) const {
    emit_partials<MB, KAHAN>(phi, kc, wgt, active, lane, dst);
  }
  __device__ __forceinline__ void store_phi(const WalkArgs& a, int m0, long long wave_base, bool active,
                                            int lane) const {
    if (!active) return;
    const long long N = a.n_walkers;
    static_for<MB>([&](auto mc) {
      constexpr int m = decltype(mc)::value;
      if (m0 + m < a.n_meas) a.phi[(size_t)(m0 + m) * N + wave_base + lane] = phi[m] - kc[m];
    });
  }
};

// Parked compensated form, for walks with many sub-steps per waveform sample:
// the phase of up to MB measurements and its Kahan compensation live in device
// memory between samples (per wavefront, groups of four measurements as
// lane-contiguous float4 rows: (group, phi | compensation, lane)), so the
// sub-step loop keeps no phase registers.  Per sample the update streams them
// through registers sixteen measurements at a time -- 2 x 16 x 4 bytes per lane
// each way, against hundreds of sub-steps of wall arithmetic -- with exactly
// PhaseVALU<MB, true>'s arithmetic per measurement, so its phases equal those of
// the register-resident compensated form bit for bit.
template <int MB>
struct PhaseParked {
  static constexpr int CH = 16;        // measurements per chunk
  float4* base;                        // this wavefront's rows
  int nch;                             // chunks holding measurements < n_meas
  __device__ __forceinline__ void init() {}
  __device__ __forceinline__ float4* row(int group, int part, int lane) const {
    return base + ((size_t)(group * 2 + part) * kWave + lane);
  }
  __device__ __forceinline__ void set_ref(const WalkArgs& a, int m0, int lane, float, float, float) {
    const size_t wave = (size_t)blockIdx.y * gridDim.x * (kBlock / kWave) +
                        (size_t)blockIdx.x * (kBlock / kWave) + (threadIdx.x / kWave);
    base = a.park + wave * (size_t)(MB / 4) * 2 * kWave;
    const int rem = a.n_meas - m0;
    nch = ((rem < MB ? rem : MB) + CH - 1) / CH;
    if (a.t_begin > 0) return;   // resumed: the rows hold the phase so far
    for (int q = 0; q < nch * (CH / 4); ++q) {
      *row(q, 0, lane) = float4{0.0f, 0.0f, 0.0f, 0.0f};
      *row(q, 1, lane) = float4{0.0f, 0.0f, 0.0f, 0.0f};
    }
  }
  __device__ __forceinline__ void prefetch(const WalkArgs&, int, int, int) {}
  __device__ __forceinline__ void add_segment(const WalkArgs&, int, int, int) {}
  __device__ __forceinline__ void load(int c, int lane, float (&phi)[CH], float (&kc)[CH]) const {
    static_for<CH / 4>([&](auto ic) {
      constexpr int i = decltype(ic)::value;
      const float4 v = *row(c * (CH / 4) + i, 0, lane), w = *row(c * (CH / 4) + i, 1, lane);
      phi[4 * i] = v.x; phi[4 * i + 1] = v.y; phi[4 * i + 2] = v.z; phi[4 * i + 3] = v.w;
      kc[4 * i] = w.x; kc[4 * i + 1] = w.y; kc[4 * i + 2] = w.z; kc[4 * i + 3] = w.w;
    });
  }
  __device__ __forceinline__ void update(const WalkArgs& a, int t, int m0, int lane, float Px, float Py, float Pz) {
#pragma unroll 1
    for (int c = 0; c < nch; ++c) {
      const cfloat* g = (const cfloat*)(a.Gs + ((size_t)t * a.n_meas_pad + m0 + c * CH) * 3);
      float phi[CH], kc[CH];
      load(c, lane, phi, kc);
      static_for<CH>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        const float dphi = fmaf(g[3 * m + 2], Pz, fmaf(g[3 * m + 1], Py, g[3 * m] * Px));
        const float yk = dphi - kc[m];
        const float tk = phi[m] + yk;
        kc[m] = (tk - phi[m]) - yk;
        phi[m] = tk;
      });
      static_for<CH / 4>([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        *row(c * (CH / 4) + i, 0, lane) = float4{phi[4 * i], phi[4 * i + 1], phi[4 * i + 2], phi[4 * i + 3]};
        *row(c * (CH / 4) + i, 1, lane) = float4{kc[4 * i], kc[4 * i + 1], kc[4 * i + 2], kc[4 * i + 3]};
      });
    }
  }
  __device__ __forceinline__ void emit(const WalkArgs&, int, float wgt, bool active, int lane, double* dst) const {
#pragma unroll 1
    for (int c = 0; c < nch; ++c) {
      float phi[CH], kc[CH];
      load(c, lane, phi, kc);
      emit_partials<CH, true>(phi, kc, wgt, active, lane, dst + 2 * CH * c);
    }
  }
  __device__ __forceinline__ void store_phi(const WalkArgs& a, int m0, long long wave_base, bool active,
                                            int lane) const {
    if (!active) return;
    const long long N = a.n_walkers;
    for (int c = 0; c < nch; ++c) {
      float phi[CH], kc[CH];
      load(c, lane, phi, kc);
      static_for<CH>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        const int mm = m0 + c * CH + m;
        if (mm < a.n_meas) a.phi[(size_t)mm * N + wave_base + lane] = phi[m] - kc[m];
      });
    }
  }
};

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef __fp16 fp16x2v __attribute__((ext_vector_type(2)));  // v_cvt_pkrtz result
typedef _Float16 half4v __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2v __attribute__((ext_vector_type(2)));

// Per-lane start point of the matrix-core phase (n_sub times the unfolded start
// position, times the caller's position factor), one slot per thread.
__device__ __forceinline__ float4* phase_ref_slots() {
  __shared__ float4 ref[kBlock];
  return ref;
}

// Waveform operands of the current sample, NP 16-byte chunks per thread, landed
// by global->LDS loads (no registers held across the sub-steps).
template <int NP>
__device__ __forceinline__ uint4* phase_b_slots() {
  __shared__ uint4 b[kBlock * NP];
  return b;
}

// Matrix-core form: per sample the wave computes the (64 walkers x 16) . (16 x MB)
// product Phi += A . B with v_mfma_f32_16x16x16f16.  With d the walker's
// displacement sum (P - n_sub * start) times p_scale and g = Gs(t) times g_scale,
// both split into float16 hi + lo parts, the K = 16 slots hold
//   A: (d_hi, 0) (d_hi, 0) (d_lo, 0) (d_lo, 0)     B: (g_hi, 0) (g_lo, 0) (g_hi, 0) (g_lo, 0)
// (each (v, 0) = x, y, z, 0), i.e. all four hi/lo products, so the product
// misses only the split residuals (relative 2^-21).  Float16 products are exact
// and accumulate in float32.  The float32 form (v_mfma_f32_16x16x4f32) runs on
// the vector ALUs' issue slots and cost four times as many cycles.  Accumulator
// tile (wt, nt) holds walkers 16 wt + 4 (lane/16) + r (r = 0..3) and measurement
// 16 nt + lane%16, in units of 1 / (p_scale g_scale); it starts at the start
// point's own phase up to the first emission point, start . sum_{t <= e_0} Gs(t),
// and after each echo emission takes the next segment's, so at every emission it
// carries the whole phase.
//
// Range: per sub-step a walker moves at most one step (three for the packed
// myelin mirror), so |d| <= 4 n_sub (n_t n_sub) step_max before scaling; the host
// picks p_scale so that bound maps to 2^14, a factor of four below the float16
// limit, and g_scale so max |Gs| maps to 2^14.
template <int MB>
struct PhaseMFMA {
  static constexpr int NT = MB / 16;
  static constexpr int NP = MB / 32;  // 16-byte operand chunks per lane: column tiles 2p, 2p + 1
  floatx4 acc[4][NT];
  __device__ __forceinline__ void init() {}
  // Records the lane's start point (n_sub x unfolded start, times the caller's
  // position factor) and starts every accumulator at that point's phase over the
  // first segment.
  __device__ __forceinline__ void set_ref(const WalkArgs& a, int m0, int lane, float rx, float ry, float rz) {
    __builtin_amdgcn_wave_barrier();
    phase_ref_slots()[threadIdx.x] = float4{rx, ry, rz, 0.0f};
    static_for<4>([&](auto wc) {
      static_for<NT>([&](auto nc) {
        acc[decltype(wc)::value][decltype(nc)::value] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
      });
    });
    add_segment(a, m0, lane, 0);
  }
  // acc += start . (sum of Gs over segment seg), in accumulator units.
  __device__ __forceinline__ void add_segment(const WalkArgs& a, int m0, int lane, int seg) {
    const int kq = lane >> 4, jq = lane & 15;
    __builtin_amdgcn_wave_barrier();
    const float4* ref = phase_ref_slots() + (threadIdx.x & ~(kWave - 1));
    const float* gs = a.Gsum + (size_t)seg * 3 * a.n_meas_pad;
    const float pg = 1.0f / a.inv_pg_scale;  // exact: a power of two
    static_for<NT>([&](auto nc) {
      constexpr int nt = decltype(nc)::value;
      // start points re-read per column tile (inside the walk loop, at echoes,
      // holding all 16 would spill)
      asm volatile("" ::: "memory");
      const int m = m0 + 16 * nt + jq;
      const float sx = gs[m], sy = gs[a.n_meas_pad + m], sz = gs[2 * a.n_meas_pad + m];
      static_for<4>([&](auto wc) {
        constexpr int wt = decltype(wc)::value;
        static_for<4>([&](auto rc) {
          constexpr int r = decltype(rc)::value;
          const float4 q = ref[16 * wt + 4 * kq + r];
          acc[wt][nt][r] = acc[wt][nt][r] + ((q.x * sx + q.y * sy) + q.z * sz) * pg;
        });
      });
    });
  }
  // B operands of sample t: lane (k = lane/16, j = lane%16) needs part k%2 (hi, lo)
  // of measurements m0 + 16 nt + j.  The table stores them as 16-byte chunks (column
  // tiles 2p and 2p + 1 of one lane), copied straight into the wave's LDS slots at
  // the start of the sample and read back in update().
  __device__ __forceinline__ void prefetch(const WalkArgs& a, int t, int m0, int lane) {
    // Uniform row base (scalar registers) plus this lane's 32-bit chunk offset.
    const uint32_t row = (uint32_t)(a.n_meas_pad / 32) * 16;
    const uint4* base = reinterpret_cast<const uint4*>(a.GsH) + ((size_t)t * 2 * row + (size_t)(m0 / 32) * 16);
    const uint4* src = base + (((uint32_t)(lane >> 4) & 1u) * row + ((uint32_t)lane & 15u));
    uint4* dst = phase_b_slots<NP>() + (threadIdx.x >> 6) * (kWave * NP);
    static_for<NP>([&](auto pc) {
      constexpr int pj = decltype(pc)::value;
      __builtin_amdgcn_global_load_lds((const void __attribute__((address_space(1)))*)(src + 16 * pj),
                                       (void __attribute__((address_space(3)))*)(dst + kWave * pj), 16, 0, 0);
    });
  }
  // A operands: every lane splits its own displacement sum and stages
  // (d_hi, 0, d_lo, 0) (16 bytes) in LDS; lane (k, j) then reads the hi (k < 2) or
  // lo (k >= 2) half of walker 16 wt + j -- four 8-byte loads per sample, no
  // bank conflicts (32 distinct 8-byte words per wave).
  __device__ __forceinline__ void update(const WalkArgs& a, int, int, int lane, float Px, float Py, float Pz) {
    __shared__ uint4 stage[kBlock];
    const int kq = lane >> 4, jq = lane & 15;
    const float4 r = phase_ref_slots()[threadIdx.x];
    const float dx = (Px - r.x) * a.p_scale, dy = (Py - r.y) * a.p_scale, dz = (Pz - r.z) * a.p_scale;
    const fp16x2v hxy = __builtin_amdgcn_cvt_pkrtz(dx, dy);
    const fp16x2v hz = __builtin_amdgcn_cvt_pkrtz(dz, 0.0f);
    const fp16x2v lxy = __builtin_amdgcn_cvt_pkrtz(dx - (float)hxy.x, dy - (float)hxy.y);
    const fp16x2v lz = __builtin_amdgcn_cvt_pkrtz(dz - (float)hz.x, 0.0f);
    __builtin_amdgcn_wave_barrier();
    stage[threadIdx.x] = make_uint4(__builtin_bit_cast(uint32_t, hxy), __builtin_bit_cast(uint32_t, hz),
                                    __builtin_bit_cast(uint32_t, lxy), __builtin_bit_cast(uint32_t, lz));
    __builtin_amdgcn_wave_barrier();
    const u32x2v* sw = reinterpret_cast<const u32x2v*>(stage + (threadIdx.x & ~(kWave - 1)));
    const uint4* bs = phase_b_slots<NP>() + (threadIdx.x >> 6) * (kWave * NP) + lane;
    half4v bv[NT];
    static_for<NP>([&](auto pc) {
      constexpr int pj = decltype(pc)::value;
      const uint4 c = bs[kWave * pj];
      bv[2 * pj] = __builtin_bit_cast(half4v, u32x2v{c.x, c.y});
      bv[2 * pj + 1] = __builtin_bit_cast(half4v, u32x2v{c.z, c.w});
    });
    static_for<4>([&](auto wc) {
      constexpr int wt = decltype(wc)::value;
      const half4v av = __builtin_bit_cast(half4v, sw[(16 * wt + jq) * 2 + (kq >> 1)]);
      static_for<NT>([&](auto nc) {
        constexpr int nt = decltype(nc)::value;
        acc[wt][nt] = __builtin_amdgcn_mfma_f32_16x16x16f16(av, bv[nt], acc[wt][nt], 0, 0, 0);
      });
    });
  }
  __device__ __forceinline__ void emit(const WalkArgs& a, int, float wgt, bool active, int lane,
                                       double* dst) const {
    const int kq = lane >> 4, jq = lane & 15;
    const float inv = a.inv_pg_scale;
    float wv[4][4];
    static_for<4>([&](auto wc) {
      constexpr int wt = decltype(wc)::value;
      static_for<4>([&](auto rc) {
        constexpr int r = decltype(rc)::value;
        const int src = 16 * wt + 4 * kq + r;
        const float ws = __shfl(wgt, src, kWave);
        const int as = __shfl(active ? 1 : 0, src, kWave);
        wv[wt][r] = as ? ws : 0.0f;
      });
    });
    static_for<NT>([&](auto nc) {
      constexpr int nt = decltype(nc)::value;
      double re = 0.0, im = 0.0;
      static_for<4>([&](auto wc) {
        constexpr int wt = decltype(wc)::value;
        static_for<4>([&](auto rc) {
          constexpr int r = decltype(rc)::value;
          float sn, cs;
          phase_sincos(acc[wt][nt][r] * inv, &sn, &cs);
          re += (double)(wv[wt][r] * cs);
          im += (double)(wv[wt][r] * sn);
        });
      });
      re += __shfl_xor(re, 16, kWave);
      im += __shfl_xor(im, 16, kWave);
      re += __shfl_xor(re, 32, kWave);
      im += __shfl_xor(im, 32, kWave);
      if (lane < 16) {
        dst[2 * (16 * nt + jq)] = re;
        dst[2 * (16 * nt + jq) + 1] = im;
      }
    });
  }
  __device__ __forceinline__ void store_phi(const WalkArgs& a, int m0, long long wave_base, bool,
                                            int lane) const {
    const int kq = lane >> 4, jq = lane & 15;
    const long long N = a.n_walkers;
    static_for<4>([&](auto wc) {
      constexpr int wt = decltype(wc)::value;
      static_for<4>([&](auto rc) {
        constexpr int r = decltype(rc)::value;
        const long long w = wave_base + 16 * wt + 4 * kq + r;
        static_for<NT>([&](auto nc) {
          constexpr int nt = decltype(nc)::value;
          const int m = m0 + 16 * nt + jq;
          if (w < N && m < a.n_meas) a.phi[(size_t)m * N + w] = acc[wt][nt][r] * a.inv_pg_scale;
        });
      });
    });
  }
};

// Kahan compensation and blocks narrower than one 16-wide matrix tile run on the
// VALU; every other block uses the matrix cores; PARK keeps the compensated phase
// in device memory between samples.
template <int MB, bool KAHAN, bool PARK = false>
using PhaseAcc = typename std::conditional<
    PARK, PhaseParked<MB>,
    typename std::conditional<(KAHAN || MB < 32), PhaseVALU<MB, KAHAN>, PhaseMFMA<MB>>::type>::type;

template <int MB, bool KAHAN>
constexpr bool uses_mfma() { return !(KAHAN || MB < 32); }

// ---------------------------------------------------------------------------
// The walk kernel.  Grid: x = walker tiles of 256, y = measurement blocks.
// ---------------------------------------------------------------------------
template <int GEOM, int WALL, int MB, bool KAHAN, bool PARK = false>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(PARK ? 4 : 2))) void walk_kernel(
    const WalkArgs a) {
  constexpr bool PACKED = (GEOM == G_PCYL || GEOM == G_PSPH);
  constexpr bool TABLES = PACKED || GEOM == G_PMYEL;
  constexpr int OBJ_ROWS = (GEOM == G_PMYEL) ? 3 : 1;
  constexpr int DIM = (GEOM == G_PSPH) ? 3 : 2;
  // The float16 trajectory channels of the walk-once producer are compiled into
  // the 4-measurement variant only (the producer launches with that block), so
  // the other variants carry no per-sub-step channel code.
  constexpr bool CHANNELS = (MB == 4 && !KAHAN);
  extern __shared__ float4 lds[];

  Tables tb{nullptr, nullptr, nullptr};
  if (TABLES) {
    const int n_cells = (DIM == 2) ? a.grid_n * a.grid_n : a.grid_n * a.grid_n * a.grid_n;
    float4* objs = lds;
    int* start = reinterpret_cast<int*>(lds + OBJ_ROWS * a.n_obj);
    int* items = start + n_cells + 1;
    for (int i = threadIdx.x; i < OBJ_ROWS * a.n_obj; i += kBlock) objs[i] = a.objs[i];
    for (int i = threadIdx.x; i <= n_cells; i += kBlock) start[i] = a.cell_start[i];
    for (int i = threadIdx.x; i < a.n_items; i += kBlock) items[i] = a.cell_items[i];
    __syncthreads();
    tb = Tables{objs, start, items};
  }

  const int lane = threadIdx.x & (kWave - 1);
  const long long gid = (long long)blockIdx.x * kBlock + threadIdx.x;
  const long long wave_id = gid / kWave;
  const bool active = gid < a.n_walkers;
  const long long N = a.n_walkers;
  const uint32_t w = (uint32_t)(active ? gid : N - 1);   // index into this launch's arrays
  const uint32_t wg = w + (uint32_t)a.first_walker;       // global walker index
  const int m0 = blockIdx.y * MB;
  const bool primary = (blockIdx.y == 0);
  const bool inject = (a.inj_n != nullptr);

  const bool resume = a.t_begin > 0;
  Pos p;
  float logw = 0.0f;
  uint32_t hits = 0, ncross = 0, ncand = 0;
  // PackedMyelinatedCylinders: the phase uses the position unwrapped by the
  // minimum-image displacement of every sub-step (physics.py:1040-1047).
  float uwx, uwy, uwz;
  if (resume) {
    const float4 s0 = a.state_in[w], s1 = a.state_in[N + w];
    const float4 s2 = (GEOM == G_PMYEL) ? a.state_in[2 * N + w] : float4{0.0f, 0.0f, 0.0f, 0.0f};
    p = Pos{s0.x, s0.y, s0.z, __float_as_int(s1.x), __float_as_int(s1.y), __float_as_int(s1.z)};
    logw = s0.w;
    ncross = __float_as_uint(s1.w);
    uwx = s2.x; uwy = s2.y; uwz = s2.z;
    if (GEOM != G_PMYEL) { uwx = p.x; uwy = p.y; uwz = p.z; }   // unused: not kept live
  } else {
    if (a.r0) {
      p = Pos{a.r0[w], a.r0[N + w], a.r0[2 * N + w], 0, 0, 0};
      if (PACKED) fold<DIM>(p, a);
      if (GEOM == G_MYEL) { p.cx = myel_classify(p.x, p.y, a); p.cy = p.cx; }
    } else {
      p = sample_start<GEOM>(wg, a, tb);
    }
    uwx = p.x; uwy = p.y; uwz = p.z;
  }
  int comp = classify<GEOM>(p, a, tb);
  if (primary && active && a.comp_origin && !resume) a.comp_origin[w] = comp;

  PhaseAcc<MB, KAHAN, PARK> ph;
  ph.init();
  const long long wave_base = wave_id * kWave;
  const float L = a.cell;
  const int n_sub = a.n_sub;
  const float fsub = (float)n_sub;
  u32x4 rnd{0u, 0u, 0u, 0u};
  int next_echo = 0;
  while (next_echo < a.n_echo && a.echo_idx[next_echo] < a.t_begin) ++next_echo;
  // Start point of the matrix-core phase: n_sub times the unfolded start position.
  if (PACKED)
    ph.set_ref(a, m0, lane, fsub * (p.x + L * (float)p.cx), fsub * (p.y + L * (float)p.cy),
               fsub * (p.z + L * (float)p.cz));
  else
    ph.set_ref(a, m0, lane, fsub * p.x, fsub * p.y, fsub * p.z);

  // Screened packed permeation (see PackedScreen).
  constexpr bool SCREEN = PACKED && WALL == W_PERMEATE;
  PackedScreen<DIM> scr;
  if (SCREEN) scr.refresh(p, a, tb);

  for (int t = a.t_begin; t < a.t_end; ++t) {
    ph.prefetch(a, t, m0, lane);
    const float chi = a.chi ? a.chi[t] : 1.0f;
    // relaxation per sub-step: T2 while transverse, T1 while stored (physics.py:307-311)
    const float q2chi = a.has_t2 ? (a.dt_sub * a.inv_T2) * chi : 0.0f;
    const float q1chi = a.has_t1 ? (a.dt_sub * a.inv_T1) * (1.0f - chi) : 0.0f;
    const float x0 = (GEOM == G_PMYEL) ? uwx : p.x;
    const float y0 = (GEOM == G_PMYEL) ? uwy : p.y;
    const float z0 = (GEOM == G_PMYEL) ? uwz : p.z;
    float ax = 0.0f, ay = 0.0f, az = 0.0f;
    float lw_sample = 0.0f, occ = 0.0f;
    for (int s = 0; s < n_sub; ++s) {
      const uint32_t gs = (uint32_t)t * (uint32_t)n_sub + (uint32_t)s;
      float ux, uy, uz, uinj = 0.0f;
      if (inject) {
        const size_t base = (size_t)gs * 3 * (size_t)N + w;
        const float nx = a.inj_n[base], ny = a.inj_n[base + N], nz = a.inj_n[base + 2 * N];
        const float nrm = sqrtf((nx * nx + ny * ny) + nz * nz);
        ux = nx / nrm; uy = ny / nrm; uz = nz / nrm;
        if ((WALL == W_PERMEATE || GEOM == G_MYEL || GEOM == G_PMYEL) && a.inj_u)
          uinj = a.inj_u[(size_t)gs * N + w];
      } else {
        if (!(gs & 1u) || (t == a.t_begin && s == 0)) {
          rnd = philox4x32_10({wg, a.instance, gs >> 1, TAG_STEP}, a.seed_lo, a.seed_hi);
        }
        const uint32_t ra = (gs & 1u) ? rnd.z : rnd.x;
        const uint32_t rb = (gs & 1u) ? rnd.w : rnd.y;
        iso_direction(ra, rb, &ux, &uy, &uz);
      }
      const PowlesSrc ps{wg, a.instance, gs, a.seed_lo, a.seed_hi, uinj, inject};
      float dlog;
      if (GEOM == G_PMYEL) {
        const float px0 = p.x, py0 = p.y, pz0 = p.z;
        const float db = pmyel_step(p, ux, uy, uz, a, tb, ps, hits, ncross, ncand);
        float dx = p.x - px0, dy = p.y - py0;
        const float dz = p.z - pz0;
        dx = dx - L * rintf(div_rn(dx, L));
        dy = dy - L * rintf(div_rn(dy, L));
        uwx = uwx + dx;
        uwy = uwy + dy;
        uwz = uwz + dz;
        // per-class T2 (transverse), T1 (stored), wall relaxivity (transverse),
        // in the reference's order (physics.py:1049-1060)
        dlog = 0.0f;
        if (a.has_t2) dlog = dlog - pick3(p.cx, a.pm_q2[0], a.pm_q2[1], a.pm_q2[2]) * chi;
        if (a.has_t1) dlog = dlog - (a.dt_sub * a.inv_T1) * (1.0f - chi);
        if (a.pm_rhoD > 0.0f) dlog = dlog + (a.pm_rhoD * db) * chi;
      } else {
        const float sl = (GEOM == G_MYEL) ? pick3(p.cx, a.m_step[0], a.m_step[1], a.m_step[2]) : a.step_len;
        const float sx = ux * sl, sy = uy * sl, sz = uz * sl;
        if (SCREEN) {
          dlog = screened_permeate<DIM>(p, sx, sy, sz, a, tb, scr, ps, hits, ncross, ncand);
        } else {
          dlog = wall_step<GEOM, WALL>(p, sx, sy, sz, a, tb, ps, hits, ncross, ncand);
        }
        // Relaxation (physics.py:307-311): walls and T2 while transverse, T1 while stored.
        if (WALL != W_REFLECT) dlog = dlog * chi;
        if (GEOM == G_MYEL) {
          // MyelinatedCylinder: dt/T2 of the compartment after the step (physics.py:657-659)
          if (a.has_t2) dlog = dlog - pick3(p.cx, a.m_q2[0], a.m_q2[1], a.m_q2[2]) * chi;
          if (a.has_t1) dlog = dlog - (a.dt_sub * a.inv_T1) * (1.0f - chi);
        } else {
          // per-sample constants, zero when the term is off (x - 0 == x, signed zeros too)
          dlog = (dlog - q2chi) - q1chi;
        }
      }
      logw = logw + dlog;
      if (CHANNELS && a.dlog_h) lw_sample = lw_sample + dlog;
      if (CHANNELS && a.comp_h) occ = occ + outside_norm<GEOM>(p, a);
      // (accumulated unconditionally: only read when n_sub > 1; a select on the
      // run-time n_sub would cost more than the sums)
      ax = ax + (((GEOM == G_PMYEL) ? uwx : p.x) - x0);
      ay = ay + (((GEOM == G_PMYEL) ? uwy : p.y) - y0);
      az = az + (((GEOM == G_PMYEL) ? uwz : p.z) - z0);
    }
    // Sum of the n_sub sub-step positions (unfolded), then the phase update.
    float Px, Py, Pz;
    {
      float ox = x0, oy = y0, oz = z0;
      float ex = (GEOM == G_PMYEL) ? uwx : p.x;
      float ey = (GEOM == G_PMYEL) ? uwy : p.y;
      float ez = (GEOM == G_PMYEL) ? uwz : p.z;
      if (PACKED) {
        ox = ox + L * (float)p.cx; oy = oy + L * (float)p.cy; oz = oz + L * (float)p.cz;
        ex = ex + L * (float)p.cx; ey = ey + L * (float)p.cy; ez = ez + L * (float)p.cz;
      }
      if (n_sub > 1) {
        Px = fmaf(fsub, ox, ax); Py = fmaf(fsub, oy, ay); Pz = fmaf(fsub, oz, az);
      } else {
        Px = ex; Py = ey; Pz = ez;
      }
    }
    ph.update(a, t, m0, lane, Px, Py, Pz);
    if (PACKED) fold<DIM>(p, a);
    if (SCREEN) scr.refresh(p, a, tb);   // the folded frame: cell, box and images anew
    if (a.comp_full) comp = classify<GEOM>(p, a, tb);
    if (primary && active) {
      if (a.comp_full) a.comp_full[(size_t)t * N + w] = comp;
      if (a.pos_full) {
        a.pos_full[((size_t)t * 3 + 0) * N + w] = PACKED ? p.x + L * (float)p.cx : p.x;
        a.pos_full[((size_t)t * 3 + 1) * N + w] = PACKED ? p.y + L * (float)p.cy : p.y;
        a.pos_full[((size_t)t * 3 + 2) * N + w] = PACKED ? p.z + L * (float)p.cz : p.z;
      }
      if (CHANNELS && a.traj_h) {
        float qx = PACKED ? p.x + L * (float)p.cx : p.x;
        float qy = PACKED ? p.y + L * (float)p.cy : p.y;
        float qz = PACKED ? p.z + L * (float)p.cz : p.z;
        if (a.traj_rotate) {
          const float* R = a.traj_R;
          const float lx = (R[0] * qx + R[1] * qy) + R[2] * qz;
          const float ly = (R[3] * qx + R[4] * qy) + R[5] * qz;
          const float lz = (R[6] * qx + R[7] * qy) + R[8] * qz;
          qx = lx; qy = ly; qz = lz;
        }
        a.traj_h[((size_t)t * 3 + 0) * N + w] = f16_bits(qx);
        a.traj_h[((size_t)t * 3 + 1) * N + w] = f16_bits(qy);
        a.traj_h[((size_t)t * 3 + 2) * N + w] = f16_bits(qz);
      }
      if (CHANNELS && a.dlog_h) a.dlog_h[(size_t)t * N + w] = f16_bits(lw_sample);
      if (CHANNELS && a.comp_h) a.comp_h[(size_t)t * N + w] = f16_bits(occ / fsub);
    }
    if (next_echo < a.n_echo && a.echo_idx[next_echo] == t) {
      float wgt = expf(logw);
      if (GEOM == G_MYEL) wgt = wgt * pick3(p.cy, a.m_wf[0], a.m_wf[1], a.m_wf[2]);
      if (GEOM == G_PMYEL && p.cz == 1) wgt = wgt * a.pm_wmy;
      ph.emit(a, m0, wgt, active, lane,
              a.echo_partials + (((size_t)wave_id * a.n_echo + next_echo) * a.n_meas_pad + m0) * 2);
      ++next_echo;
      ph.add_segment(a, m0, lane, next_echo);
    }
  }

  if (a.t_end < a.n_t) {
    // more samples follow in another launch: keep the walker, not the ensemble
    if (primary && active) {
      a.state_out[w] = float4{p.x, p.y, p.z, logw};
      a.state_out[N + w] = float4{__int_as_float(p.cx), __int_as_float(p.cy), __int_as_float(p.cz),
                                  __uint_as_float(ncross)};
      if (GEOM == G_PMYEL) a.state_out[2 * N + w] = float4{uwx, uwy, uwz, 0.0f};
    }
  } else {
    if (a.comp_final && !a.comp_full) comp = classify<GEOM>(p, a, tb);
    float wgt = expf(logw);
    // MyelinatedCylinder: water-fraction spin weight of the start compartment (_ens, core.py:426-429, 489)
    if (GEOM == G_MYEL) wgt = wgt * pick3(p.cy, a.m_wf[0], a.m_wf[1], a.m_wf[2]);
    // PackedMyelinatedCylinders: myelin water carries its proton density (core.py:552-553)
    if (GEOM == G_PMYEL && p.cz == 1) wgt = wgt * a.pm_wmy;
    ph.emit(a, m0, wgt, active, lane, a.partials + ((size_t)wave_id * a.n_meas_pad + m0) * 2);
    if (a.phi) ph.store_phi(a, m0, wave_base, active, lane);
    if (active && primary) {
      if (a.final_pos) {
        a.final_pos[w] = PACKED ? p.x + L * (float)p.cx : p.x;
        a.final_pos[N + w] = PACKED ? p.y + L * (float)p.cy : p.y;
        a.final_pos[2 * N + w] = PACKED ? p.z + L * (float)p.cz : p.z;
      }
      if (a.comp_final) a.comp_final[w] = comp;
      if (a.n_cross) a.n_cross[w] = ncross;
      if (a.log_w) a.log_w[w] = logw;
    }
  }
  if (a.counters && primary) {
    unsigned long long h = active ? hits : 0u, c = active ? ncand : 0u;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      h += __shfl_xor(h, off, kWave);
      c += __shfl_xor(c, off, kWave);
    }
    if (lane == 0) {
      atomicAdd(&a.counters[0], h);
      atomicAdd(&a.counters[1], c);
    }
  }
}

}  // namespace dmw
