/*
 * walk_oracle.c -- CPU restatement of dmipy-sim's fused Monte-Carlo walk.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker of the HIP walk: it is
 * built and called by tests/, __graft_entry__.smoke() and the cpu_baseline leg
 * of bench.py, never by the dmipy_sim_amd package.
 *
 * It restates, operation by operation in float32, the reference's step:
 *   sub-step body          dmipy_sim/physics.py:296-314 (permeable),
 *                          337-355 (surface), 371-387 (T2/T1), 402-412 (plain)
 *   FreeDiffusion          dmipy_sim/geometries.py:40-41
 *   Box1D reflect(_w)      dmipy_sim/geometries.py:93-139
 *   Sphere reflect(_w)     dmipy_sim/geometries.py:197-317
 *   Sphere permeate        dmipy_sim/geometries.py:320-424
 *   Cylinder reflect(_w)   dmipy_sim/geometries.py:516-654
 *   Cylinder permeate      dmipy_sim/geometries.py:656-801
 *   classify_position      dmipy_sim/geometries.py:180-184, 803-813, 2185-2214, 2758-2771
 *   PackedCyl. reflect(_w) dmipy_sim/geometries.py:1746-2016
 *   PackedCyl. permeate    dmipy_sim/geometries.py:2018-2183
 *   PackedSph. reflect(_w) dmipy_sim/geometries.py:2445-2631
 *   PackedSph. permeate    dmipy_sim/geometries.py:2633-2756
 *   ensemble               dmipy_sim/core.py:426-429
 * with the reference's cost structure: the fixed 10-iteration bounce scan,
 * every object of a packing tested on every sub-step (plus the all-object
 * safety-clamp pass), and the phase of every measurement advanced on every
 * sub-step with the float32 right-endpoint rule phi += gamma*dt_sub*(g.r_new).
 *
 * Choices the reference leaves open (its XLA lowering decides them) are fixed
 * here: 3-term sums are ((a+b)+c), the per-bounce d_perp values are summed in
 * iteration order, no contraction into FMA (build with -ffp-contract=off).
 *
 * Random numbers: the reference's threefry stream cannot be reproduced without
 * JAX; the oracle consumes either injected raw normal triples / Powles uniforms
 * (the reference's increments), or a Philox4x32-10 stream (Salmon et al., SC'11)
 * with the isotropic-direction construction documented in oracle_direction().
 *
 * Position forms for packed substrates: 0 = unfolded float32 coordinates, as
 * the reference keeps them (geometries.py:1843-1849); 1 = integer cell index +
 * float32 offset folded to the primary cell after every waveform sample.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
  int32_t kind, wall, n_obj;
  float size, cell, eps, nudge, rho_over_D, kappa_over_D;
  const float* centers;
  const float* radii;
  int32_t n_sub;
  float step_len, gamma_dt_sub, dt_sub, inv_T2, inv_T1;
  int32_t has_t2, has_t1;
  int32_t n_t, n_meas;
  const float* G;          /* (n_meas, n_t, 3) */
  const float* chi;        /* (n_t) or NULL */
  int32_t n_echo;
  const int32_t* echo_idx; /* (n_echo) */
  int64_t n_walkers;
  uint64_t seed;
  const float* r0;         /* (N, 3) or NULL: Philox start sampler */
  const float* inj_n;      /* (N, S, 3) or NULL */
  const float* inj_u;      /* (N, S) or NULL */
  int32_t cell_form;
  int32_t phase64;         /* 1: phase accumulated in float64 instead of float32 */
  int32_t threads;
  int64_t walker_begin;    /* sub-range of walkers to run (cpu timing slices) */
  int64_t walker_end;      /* 0 = all */
  uint32_t instance;       /* second Philox counter word (substrate instance) */
  float axes[3];           /* Ellipsoid: float32 1/a^2, 1/b^2, 1/c^2 */
  float aux;               /* Ellipsoid: 1 - NUDGE/min semi-axis; shell, myelin: R_out */
  /* MyelinatedCylinder, per compartment (intra, sheath, extra) */
  float m_step[3];         /* float32 sqrt(6 D_c dt) */
  float m_q2[3];           /* float32 dt / T2_c */
  float m_wf[3];           /* spin weights (water fractions) */
  float m_k2[4];           /* 2 kappa / max(D_leaving, 1e-30): inner from 0, 1; outer from 1, 2 */
  int64_t m_n0, m_n01;     /* start compartments by walker index */
  /* PackedMyelinatedCylinders: centers (n_obj, 2), radii = outer radii */
  const float* pm_inner;   /* (n_obj) inner radii */
  const float* pm_step;    /* (n_obj, 3) step lengths intra, myelin, extra */
  const float* pm_k2;      /* (n_obj, 4) 2 kappa/D: inner from intra, from myelin; outer from myelin, from extra */
  float pm_q2[3];          /* dt_sub / T2 per class: intra, myelin, extra */
  float pm_rhoD;           /* wall relaxivity / D_ref */
  float pm_wmy;            /* myelin spin weight */
  const int64_t* pm_start; /* (2 n_obj + 1) start boundaries */
  int64_t first_walker;    /* global index of walker 0 (keys the Philox streams and starts) */
} oracle_args;

typedef struct {
  double* signal;          /* (rows, n_meas) */
  double* signal_im;
  float* final_pos;        /* (N, 3) unfolded */
  int32_t* comp_origin;
  int32_t* comp_final;
  int32_t* comp_full;      /* (N, n_t) */
  uint32_t* n_cross;
  float* log_w;
  float* phi;              /* (N, n_meas) */
  uint64_t n_hits;
  float* pos_full;         /* (N, n_t, 3) unfolded position after every sample, or NULL */
  float* dlog_full;        /* (N, n_t) log-weight summed over each sample's sub-steps, or NULL */
  float* occ_full;         /* (N, n_t) fraction of each sample's sub-steps outside, or NULL */
  /* The exact phase of the walk's float32 trajectory, alongside (NULL = skip): each
     increment gamma_dt_sub * (g . r) formed and summed in float64 from the float32
     positions (unfolded in float64 for packed substrates). */
  double* signal_x;        /* (rows, n_meas) Re <w e^{i phi_exact}> */
  double* signal_x_im;
  double* phi_x;           /* (N, n_meas) */
} oracle_out;

enum { K_FREE = 0, K_BOX = 1, K_CYL = 2, K_SPH = 3, K_PCYL = 4, K_PSPH = 5,
       K_ELL = 6, K_SLAB = 7, K_SHELL = 8, K_SHELLC = 9, K_MYEL = 10, K_PMYEL = 11 };
enum { W_REFLECT = 0, W_SURFACE = 1, W_PERMEATE = 2 };

/* ---------------------------------------------------------------- Philox */
static void philox10(const uint32_t in[4], uint32_t k0, uint32_t k1, uint32_t out[4]) {
  uint32_t v[4] = {in[0], in[1], in[2], in[3]};
  for (int round = 0; round < 10; ++round) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * v[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * v[2];
    const uint32_t a = (uint32_t)(p1 >> 32) ^ v[1] ^ k0;
    const uint32_t b = (uint32_t)p1;
    const uint32_t c = (uint32_t)(p0 >> 32) ^ v[3] ^ k1;
    const uint32_t d = (uint32_t)p0;
    v[0] = a; v[1] = b; v[2] = c; v[3] = d;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  memcpy(out, v, sizeof v);
}

void oracle_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  philox10(ctr, key[0], key[1], out);
}

static float unit_u(uint32_t b) { return (float)(b >> 8) * (1.0f / 16777216.0f); }

/* Isotropic unit vector from two 32-bit words: z = 1 - 2u (uniform on (-1,1]),
 * azimuth = 2*pi*v; sin/cos(pi*f) on |f| <= 1/4 by their Taylor polynomials
 * (degree 9 / 10, float32 coefficients (-1)^k pi^n / n!), Horner with fmaf. */
static void oracle_direction(uint32_t a, uint32_t b, float d[3]) {
  const float z = 1.0f - 2.0f * unit_u(a);
  const float s = sqrtf(fmaxf(1.0f - z * z, 0.0f));
  const float t = 2.0f * unit_u(b);
  const float q = rintf(2.0f * t);
  const float f = t - 0.5f * q;
  const float x = f * f;
  float ps = fmaf(x, 0.08214588463306427f, -0.5992645025253296f);
  ps = fmaf(x, ps, 2.550163984298706f);
  ps = fmaf(x, ps, -5.167712688446045f);
  ps = fmaf(x, ps, 3.1415927410125732f);
  const float sf = f * ps;
  float pc = fmaf(x, -0.025806890800595284f, 0.23533062636852264f);
  pc = fmaf(x, pc, -1.3352627754211426f);
  pc = fmaf(x, pc, 4.058712005615234f);
  pc = fmaf(x, pc, -4.934802055358887f);
  const float cf = fmaf(x, pc, 1.0f);
  float c, sn;
  switch (((int)q) & 3) {
    case 0: c = cf; sn = sf; break;
    case 1: c = -sf; sn = cf; break;
    case 2: c = -cf; sn = -sf; break;
    default: c = sf; sn = -cf; break;
  }
  d[0] = s * c;
  d[1] = s * sn;
  d[2] = z;
}

/* ------------------------------------------------------------- helpers */
static float dot3(const float a[3], const float b[3]) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }
static float norm3(const float a[3]) { return sqrtf(dot3(a, a)); }

/* x - L*floor(x/L + 0.5) component (min-image / fold), reference form. */
static float mi(float q, float L) { return q - L * floorf(q / L + 0.5f); }

/* -------------------------------------------------------------- Box1D */
static float box_step(float r[3], const float st[3], float L, float rhoD, int surface, uint32_t* hits) {
  const float xr = r[0] + st[0];
  float dlog = 0.0f;
  if (surface) {
    const float dperp = fmaxf(xr - L, 0.0f) + fmaxf(-xr, 0.0f);
    if (dperp > 0.0f) { dlog = -2.0f * rhoD * dperp; (*hits)++; }
  }
  const float L2 = 2.0f * L;
  float m = fmodf(xr, L2);                      /* jnp.mod = lax.rem + sign fix */
  if (m != 0.0f && ((m < 0.0f) != (L2 < 0.0f))) m = m + L2;
  if (m > L) m = L2 - m;
  r[0] = m;
  r[1] = r[1] + st[1];
  r[2] = r[2] + st[2];
  return dlog;
}

/* ---------------------------------------------- Cylinder (interior, xy) */
static float cyl_reflect(float r[3], const float st[3], const oracle_args* A, uint32_t* hits) {
  const float R = A->size, EPS = A->eps, NUDGE = A->nudge;
  const float sxy[2] = {st[0], st[1]};
  const float sl = sqrtf(sxy[0] * sxy[0] + sxy[1] * sxy[1]);
  float d2[2] = {0.0f, 0.0f};
  if (sl > 0.0f) { d2[0] = sxy[0] / sl; d2[1] = sxy[1] / sl; }
  float p2[2] = {r[0], r[1]};
  float rem = sl, dsum = 0.0f;
  for (int it = 0; it < 10; ++it) {             /* fixed-length scan */
    const float dp = d2[0] * p2[0] + d2[1] * p2[1];
    const float disc = fmaxf(dp * dp - ((p2[0] * p2[0] + p2[1] * p2[1]) - R * R), 0.0f);
    const float d = -dp + sqrtf(disc);
    const int hit = (d > EPS) && (d < rem) && (sl > 0.0f);
    const float h[2] = {p2[0] + d * d2[0], p2[1] + d * d2[1]};
    const float n[2] = {h[0] / R, h[1] / R};
    const float dd = 2.0f * (d2[0] * n[0] + d2[1] * n[1]);
    float rf[2] = {d2[0] - dd * n[0], d2[1] - dd * n[1]};
    const float rn = sqrtf(rf[0] * rf[0] + rf[1] * rf[1]);
    rf[0] = rf[0] / rn; rf[1] = rf[1] / rn;
    const float cosa = sqrtf(disc) / R;
    const float dperp = hit ? (rem - d) * cosa : 0.0f;
    if (hit) {
      p2[0] = h[0] - NUDGE * n[0]; p2[1] = h[1] - NUDGE * n[1];
      d2[0] = rf[0]; d2[1] = rf[1];
      rem = rem - d - NUDGE;
      (*hits)++;
    }
    dsum = dsum + dperp;
  }
  const float rm = fmaxf(rem, 0.0f);
  float xy[2] = {p2[0] + d2[0] * rm, p2[1] + d2[1] * rm};
  const float nn = sqrtf(xy[0] * xy[0] + xy[1] * xy[1]);
  if (nn >= R) { xy[0] = xy[0] * (R - NUDGE) / nn; xy[1] = xy[1] * (R - NUDGE) / nn; }
  r[0] = xy[0]; r[1] = xy[1]; r[2] = r[2] + st[2];
  return dsum;
}

static float cyl_permeate(float r[3], const float st[3], float u, const oracle_args* A,
                          uint32_t* hits, uint32_t* ncross) {
  const float R = A->size, EPS = A->eps, NUDGE = A->nudge;
  const float p2[2] = {r[0], r[1]};
  const float sxy[2] = {st[0], st[1]};
  const float sl = sqrtf(sxy[0] * sxy[0] + sxy[1] * sxy[1]);
  float dh[2] = {0.0f, 0.0f};
  if (sl > 0.0f) { dh[0] = sxy[0] / sl; dh[1] = sxy[1] / sl; }
  const float dp = dh[0] * p2[0] + dh[1] * p2[1];
  const float disc = dp * dp - ((p2[0] * p2[0] + p2[1] * p2[1]) - R * R);
  const float disc_s = fmaxf(disc, 0.0f);
  const int inside = (p2[0] * p2[0] + p2[1] * p2[1]) < R * R;
  const float t_exit = -dp + sqrtf(disc_s), t_entry = -dp - sqrtf(disc_s);
  const float t_hit = inside ? t_exit : t_entry;
  const int any_hit = (disc > 0.0f) && (t_hit > EPS) && (t_hit < sl) && (sl > 0.0f);
  const float ts = any_hit ? t_hit : 0.0f;
  const float hr[2] = {p2[0] + ts * dh[0], p2[1] + ts * dh[1]};
  const float hl = sqrtf(hr[0] * hr[0] + hr[1] * hr[1]);
  const float n[2] = {hr[0] / fmaxf(hl, 1e-30f), hr[1] / fmaxf(hl, 1e-30f)};
  const float hs[2] = {R * n[0], R * n[1]};
  const float remaining = sl - ts;
  const float cosa = sqrtf(disc_s) / R;
  const float dperp = any_hit ? remaining * cosa : 0.0f;
  const float pt = fminf(1.0f, 2.0f * A->kappa_over_D * dperp);
  const int transmit = any_hit && (u < pt);
  const float straight[2] = {p2[0] + sxy[0], p2[1] + sxy[1]};
  const float dd = 2.0f * (dh[0] * n[0] + dh[1] * n[1]);
  float rf[2] = {dh[0] - dd * n[0], dh[1] - dd * n[1]};
  const float rn = sqrtf(rf[0] * rf[0] + rf[1] * rf[1]);
  rf[0] = rf[0] / rn; rf[1] = rf[1] / rn;
  const float nd[2] = {inside ? -n[0] : n[0], inside ? -n[1] : n[1]};
  const float nud[2] = {hs[0] + NUDGE * nd[0], hs[1] + NUDGE * nd[1]};
  const float go = fmaxf(remaining - NUDGE, 0.0f);
  float rr[2] = {nud[0] + rf[0] * go, nud[1] + rf[1] * go};
  const float rrn = sqrtf(rr[0] * rr[0] + rr[1] * rr[1]);
  const float rrs = fmaxf(rrn, 1e-30f);
  const float target = inside ? R - NUDGE : R + NUDGE;
  const int wrong = inside ? (rrn >= R) : (rrn <= R);
  if (wrong) { rr[0] = rr[0] * target / rrs; rr[1] = rr[1] * target / rrs; }
  float xy[2];
  if (any_hit) {
    xy[0] = transmit ? straight[0] : rr[0];
    xy[1] = transmit ? straight[1] : rr[1];
    (*hits)++;
    if (transmit) (*ncross)++;
  } else {
    xy[0] = p2[0] + sxy[0]; xy[1] = p2[1] + sxy[1];
  }
  r[0] = xy[0]; r[1] = xy[1]; r[2] = r[2] + st[2];
  return (any_hit && !transmit) ? -2.0f * A->rho_over_D * dperp : 0.0f;
}

/* ------------------------------------------------ Sphere (interior, 3D) */
static float sph_reflect(float r[3], const float st[3], const oracle_args* A, uint32_t* hits) {
  const float R = A->size, EPS = A->eps, NUDGE = A->nudge;
  const float sl = norm3(st);
  float dh[3] = {st[0] / sl, st[1] / sl, st[2] / sl};
  float p[3] = {r[0], r[1], r[2]};
  float rem = sl, dsum = 0.0f;
  for (int it = 0; it < 10; ++it) {
    const float dp = dot3(dh, p);
    const float disc = fmaxf(dp * dp - (dot3(p, p) - R * R), 0.0f);
    const float d = -dp + sqrtf(disc);
    const int hit = (d > EPS) && (d < rem);
    const float h[3] = {p[0] + d * dh[0], p[1] + d * dh[1], p[2] + d * dh[2]};
    const float n[3] = {h[0] / R, h[1] / R, h[2] / R};
    const float dd = 2.0f * dot3(dh, n);
    float rf[3] = {dh[0] - dd * n[0], dh[1] - dd * n[1], dh[2] - dd * n[2]};
    const float rn = norm3(rf);
    rf[0] = rf[0] / rn; rf[1] = rf[1] / rn; rf[2] = rf[2] / rn;
    const float dperp = hit ? (rem - d) * (sqrtf(disc) / R) : 0.0f;
    if (hit) {
      for (int c = 0; c < 3; ++c) { p[c] = h[c] - NUDGE * n[c]; dh[c] = rf[c]; }
      rem = rem - d - NUDGE;
      (*hits)++;
    }
    dsum = dsum + dperp;
  }
  const float rm = fmaxf(rem, 0.0f);
  float o[3] = {p[0] + dh[0] * rm, p[1] + dh[1] * rm, p[2] + dh[2] * rm};
  const float nn = norm3(o);
  if (nn >= R) for (int c = 0; c < 3; ++c) o[c] = o[c] * (R - NUDGE) / nn;
  r[0] = o[0]; r[1] = o[1]; r[2] = o[2];
  return dsum;
}

static float sph_permeate(float r[3], const float st[3], float u, const oracle_args* A,
                          uint32_t* hits, uint32_t* ncross) {
  const float R = A->size, EPS = A->eps, NUDGE = A->nudge;
  const float sl = norm3(st);
  const float dh[3] = {st[0] / sl, st[1] / sl, st[2] / sl};
  const float dp = dot3(dh, r);
  const float disc = dp * dp - (dot3(r, r) - R * R);
  const float disc_s = fmaxf(disc, 0.0f);
  const int inside = dot3(r, r) < R * R;
  const float t_hit = inside ? (-dp + sqrtf(disc_s)) : (-dp - sqrtf(disc_s));
  const int any_hit = (disc > 0.0f) && (t_hit > EPS) && (t_hit < sl) && (sl > 0.0f);
  const float ts = any_hit ? t_hit : 0.0f;
  const float h[3] = {r[0] + ts * dh[0], r[1] + ts * dh[1], r[2] + ts * dh[2]};
  const float n[3] = {h[0] / R, h[1] / R, h[2] / R};
  const float remaining = sl - ts;
  const float dperp = any_hit ? remaining * (sqrtf(disc_s) / R) : 0.0f;
  const float pt = fminf(1.0f, 2.0f * A->kappa_over_D * dperp);
  const int transmit = any_hit && (u < pt);
  const float dd = 2.0f * dot3(dh, n);
  float rf[3] = {dh[0] - dd * n[0], dh[1] - dd * n[1], dh[2] - dd * n[2]};
  const float rn = norm3(rf);
  for (int c = 0; c < 3; ++c) rf[c] = rf[c] / rn;
  const float go = fmaxf(remaining - NUDGE, 0.0f);
  float out[3];
  for (int c = 0; c < 3; ++c) {
    const float nd = inside ? -n[c] : n[c];
    const float refl = (h[c] + NUDGE * nd) + rf[c] * go;
    const float straight = r[c] + st[c];
    out[c] = any_hit ? (transmit ? straight : refl) : r[c] + st[c];
  }
  if (any_hit) { (*hits)++; if (transmit) (*ncross)++; }
  r[0] = out[0]; r[1] = out[1]; r[2] = out[2];
  return (any_hit && !transmit) ? -2.0f * A->rho_over_D * dperp : 0.0f;
}

/* ------------------------------------- packed substrates (brute force) */
static void obj_center(const oracle_args* A, int k, int dim, float c[3]) {
  c[0] = A->centers[k * dim];
  c[1] = A->centers[k * dim + 1];
  c[2] = dim == 3 ? A->centers[k * dim + 2] : 0.0f;
}

static float dotd(const float a[3], const float b[3], int dim) {
  return dim == 2 ? a[0] * b[0] + a[1] * b[1] : dot3(a, b);
}

/* PackedCylinders.reflect / reflect_with_log_weight, PackedSpheres likewise.
 * Returns d_perp of the hit (0 without hit). */
static float packed_reflect(float r[3], const float st[3], const oracle_args* A, int dim, uint32_t* hits) {
  const float L = A->cell, EPS = A->eps, NUDGE = A->nudge;
  const float s[3] = {st[0], st[1], dim == 3 ? st[2] : 0.0f};
  const float sl = sqrtf(dotd(s, s, dim));
  float dh[3] = {0.0f, 0.0f, 0.0f};
  if (sl > 0.0f) for (int c = 0; c < dim; ++c) dh[c] = s[c] / sl;
  int imin = 0;
  float tmin = INFINITY, disc_min = 0.0f;
  for (int k = 0; k < A->n_obj; ++k) {          /* all objects, argmin first index */
    float c[3], q[3] = {0, 0, 0};
    obj_center(A, k, dim, c);
    for (int d = 0; d < dim; ++d) q[d] = mi(r[d] - c[d], L);
    const float Rk = A->radii[k];
    const float dp = dotd(dh, q, dim);
    const float disc = dp * dp - (dotd(q, q, dim) - Rk * Rk);
    const float t = -dp - sqrtf(fmaxf(disc, 0.0f));
    const int valid = (disc > 0.0f) && (t > EPS) && (t < sl) && (sl > 0.0f);
    const float tv = valid ? t : INFINITY;
    if (tv < tmin) { tmin = tv; imin = k; disc_min = disc; }
  }
  const int any_hit = isfinite(tmin);
  const float ts = any_hit ? tmin : 0.0f;
  float c[3], qc[3] = {0, 0, 0};
  obj_center(A, imin, dim, c);
  const float Rh = A->radii[imin];
  for (int d = 0; d < dim; ++d) qc[d] = mi(r[d] - c[d], L);
  float n[3] = {0, 0, 0};
  for (int d = 0; d < dim; ++d) n[d] = (qc[d] + ts * dh[d]) / Rh;
  const float dd = 2.0f * dotd(dh, n, dim);
  float rf[3] = {0, 0, 0};
  for (int d = 0; d < dim; ++d) rf[d] = dh[d] - dd * n[d];
  const float rfn = sqrtf(dotd(rf, rf, dim));
  for (int d = 0; d < dim; ++d) rf[d] = rfn > 0.0f ? rf[d] / fmaxf(rfn, 1e-30f) : 0.0f;
  const float remaining = sl - ts;
  const float go = fmaxf(remaining - NUDGE, 0.0f);
  float f[3];
  for (int d = 0; d < dim; ++d) {
    const float refl = ((r[d] + ts * dh[d]) + NUDGE * n[d]) + rf[d] * go;
    f[d] = any_hit ? refl : r[d] + s[d];
  }
  /* safety clamp over all objects on the folded end point */
  float fo[3] = {0, 0, 0};
  for (int d = 0; d < dim; ++d) fo[d] = mi(f[d], L);
  int kcl = 0;
  float pmin = INFINITY, qcl[3] = {0, 0, 0};
  for (int k = 0; k < A->n_obj; ++k) {
    float ck[3], q[3] = {0, 0, 0};
    obj_center(A, k, dim, ck);
    for (int d = 0; d < dim; ++d) q[d] = mi(fo[d] - ck[d], L);
    const float d2 = dotd(q, q, dim);
    const float Rk = A->radii[k];
    const float pen = d2 < Rk * Rk ? d2 / (Rk * Rk) : 1.0f;
    if (pen < pmin) { pmin = pen; kcl = k; memcpy(qcl, q, sizeof q); }
  }
  const int inside_any = pmin < 1.0f;
  const float dcl = sqrtf(dotd(qcl, qcl, dim));
  const float scale = inside_any ? (A->radii[kcl] + NUDGE) / fmaxf(dcl, NUDGE) : 1.0f;
  for (int d = 0; d < dim; ++d) f[d] = f[d] + qcl[d] * (scale - 1.0f);
  r[0] = f[0]; r[1] = f[1];
  r[2] = dim == 3 ? f[2] : r[2] + st[2];
  if (any_hit) (*hits)++;
  return any_hit ? remaining * (sqrtf(fmaxf(disc_min, 0.0f)) / Rh) : 0.0f;
}

/* The per-object root test of PackedCylinders/PackedSpheres.permeate
   (geometries.py:2067-2093, 2667-2691): minimum-image offset, inside flag, the
   exit root from inside or the entry root from outside; valid in (EPS, |s|). */
static int permeate_root(const float r[3], const float c[3], float Rk, const float dh[3], float sl,
                         float L, float EPS, int dim, float* t_out, float* disc_out, int* inside_out) {
  float q[3] = {0, 0, 0};
  for (int d = 0; d < dim; ++d) q[d] = mi(r[d] - c[d], L);
  const float dist2 = dotd(q, q, dim);
  const int inside = dist2 < Rk * Rk;
  const float dp = dotd(dh, q, dim);
  const float disc = dp * dp - (dist2 - Rk * Rk);
  const float sq = sqrtf(fmaxf(disc, 0.0f));
  const float t = inside ? (-dp + sq) : (-dp - sq);
  *t_out = t;
  *disc_out = disc;
  *inside_out = inside;
  return (disc > 0.0f) && (t > EPS) && (t < sl) && (sl > 0.0f);
}

/* Batch of single-object root tests (for checking the GPU walk's wall screen):
   walker r[i], step st[i] (its first dim components form the geometry's step, as in
   packed_permeate), object centre c[i] and radius R[i]; valid[i] = 1 when the
   reference finds a valid root. */
void oracle_permeate_roots(int64_t n, const float* r, const float* st, const float* c,
                           const float* R, float L, float eps, int32_t dim, int32_t* valid) {
  for (int64_t i = 0; i < n; ++i) {
    const float s[3] = {st[3 * i], st[3 * i + 1], dim == 3 ? st[3 * i + 2] : 0.0f};
    const float sl = sqrtf(dotd(s, s, dim));
    float dh[3] = {0.0f, 0.0f, 0.0f};
    if (sl > 0.0f) for (int d = 0; d < dim; ++d) dh[d] = s[d] / sl;
    float t, disc;
    int inside;
    valid[i] = permeate_root(r + 3 * i, c + 3 * i, R[i], dh, sl, L, eps, dim, &t, &disc, &inside);
  }
}

static float packed_permeate(float r[3], const float st[3], float u, const oracle_args* A, int dim,
                             uint32_t* hits, uint32_t* ncross) {
  const float L = A->cell, EPS = A->eps, NUDGE = A->nudge;
  const float s[3] = {st[0], st[1], dim == 3 ? st[2] : 0.0f};
  const float sl = sqrtf(dotd(s, s, dim));
  float dh[3] = {0.0f, 0.0f, 0.0f};
  if (sl > 0.0f) for (int c = 0; c < dim; ++c) dh[c] = s[c] / sl;
  int imin = 0, in_min = 0;
  float tmin = INFINITY, disc_min = 0.0f;
  for (int k = 0; k < A->n_obj; ++k) {
    float c[3], t, disc;
    int inside;
    obj_center(A, k, dim, c);
    const int valid = permeate_root(r, c, A->radii[k], dh, sl, L, EPS, dim, &t, &disc, &inside);
    const float tv = valid ? t : INFINITY;
    if (tv < tmin) { tmin = tv; imin = k; disc_min = disc; in_min = inside; }
  }
  const int any_hit = isfinite(tmin);
  if (!any_hit) {
    r[0] = r[0] + st[0]; r[1] = r[1] + st[1]; r[2] = r[2] + st[2];
    return 0.0f;
  }
  const float ts = tmin;
  float c[3], qc[3] = {0, 0, 0};
  obj_center(A, imin, dim, c);
  const float Rh = A->radii[imin];
  for (int d = 0; d < dim; ++d) qc[d] = mi(r[d] - c[d], L);
  float hr[3] = {0, 0, 0};
  for (int d = 0; d < dim; ++d) hr[d] = qc[d] + ts * dh[d];
  const float hl = sqrtf(dotd(hr, hr, dim));
  float n[3] = {0, 0, 0}, hs[3] = {0, 0, 0};
  for (int d = 0; d < dim; ++d) { n[d] = hr[d] / fmaxf(hl, 1e-30f); hs[d] = Rh * n[d]; }
  const float remaining = sl - ts;
  const float dperp = remaining * (sqrtf(fmaxf(disc_min, 0.0f)) / Rh);
  const float pt = fminf(1.0f, 2.0f * A->kappa_over_D * dperp);
  const int transmit = u < pt;
  (*hits)++;
  if (transmit) {
    (*ncross)++;
    r[0] = r[0] + st[0]; r[1] = r[1] + st[1]; r[2] = r[2] + st[2];
    return 0.0f;
  }
  const float dd = 2.0f * dotd(dh, n, dim);
  float rf[3] = {0, 0, 0};
  for (int d = 0; d < dim; ++d) rf[d] = dh[d] - dd * n[d];
  const float rfn = sqrtf(dotd(rf, rf, dim));
  for (int d = 0; d < dim; ++d) rf[d] = rfn > 0.0f ? rf[d] / fmaxf(rfn, 1e-30f) : 0.0f;
  const float go = fmaxf(remaining - NUDGE, 0.0f);
  float qr[3] = {0, 0, 0};
  for (int d = 0; d < dim; ++d) {
    const float nd = in_min ? -n[d] : n[d];
    qr[d] = (hs[d] + NUDGE * nd) + rf[d] * go;
  }
  const float qn = sqrtf(dotd(qr, qr, dim));
  const float qs = fmaxf(qn, 1e-30f);
  const float target = in_min ? Rh - NUDGE : Rh + NUDGE;
  const int wrong = in_min ? (qn >= Rh) : (qn <= Rh);
  if (wrong) for (int d = 0; d < dim; ++d) qr[d] = qr[d] * target / qs;
  for (int d = 0; d < dim; ++d) r[d] = r[d] + (qr[d] - qc[d]);
  if (dim == 2) r[2] = r[2] + st[2];
  return -2.0f * A->rho_over_D * dperp;
}

/* ---------------------------------------------- Ellipsoid (geometries.py:1137-1425) */
/* u . diag(1/a^2, 1/b^2, 1/c^2) . v, as jnp.dot(u * inv_semi_sq, v) */
static float ell_q(const float u[3], const float v[3], const float inv[3]) {
  const float w[3] = {u[0] * inv[0], u[1] * inv[1], u[2] * inv[2]};
  return dot3(w, v);
}

/* reflect / reflect_with_log_weight: 10-iteration scan, then the clamp in
   normalised coordinates (geometries.py:1187-1300); returns sum d_perp. */
static float ell_reflect(float r[3], const float st[3], const oracle_args* A, uint32_t* hits) {
  const float* inv = A->axes;
  const float EPS = A->eps, NUDGE = A->nudge;
  const float sl = norm3(st);
  float d[3] = {st[0] / sl, st[1] / sl, st[2] / sl};
  float p[3] = {r[0], r[1], r[2]};
  float rem = sl, dsum = 0.0f;
  for (int it = 0; it < 10; ++it) {
    const float Aq = ell_q(d, d, inv), Bq = ell_q(p, d, inv), Cq = ell_q(p, p, inv) - 1.0f;
    const float disc = fmaxf(Bq * Bq - Aq * Cq, 0.0f);
    const float t = (-Bq + sqrtf(disc)) / Aq;
    if (!(t > EPS && t < rem)) continue;   /* the scan's no-op iteration */
    float h[3], n[3], rf[3];
    for (int k = 0; k < 3; ++k) h[k] = p[k] + t * d[k];
    const float m[3] = {h[0] * inv[0], h[1] * inv[1], h[2] * inv[2]};
    const float mn = norm3(m);
    for (int k = 0; k < 3; ++k) n[k] = m[k] / mn;
    const float ca = dot3(d, n);
    for (int k = 0; k < 3; ++k) rf[k] = d[k] - (2.0f * ca) * n[k];
    const float rn = norm3(rf);
    for (int k = 0; k < 3; ++k) rf[k] = rf[k] / rn;
    dsum = dsum + (rem - t) * ca;
    for (int k = 0; k < 3; ++k) { p[k] = h[k] - NUDGE * n[k]; d[k] = rf[k]; }
    rem = (rem - t) - NUDGE;
    (*hits)++;
  }
  const float rm = fmaxf(rem, 0.0f);
  float f[3];
  for (int k = 0; k < 3; ++k) f[k] = p[k] + d[k] * rm;
  const float qv[3] = {f[0] * sqrtf(inv[0]), f[1] * sqrtf(inv[1]), f[2] * sqrtf(inv[2])};
  const float q = norm3(qv);
  if (q >= 1.0f)
    for (int k = 0; k < 3; ++k) f[k] = (f[k] * A->aux) / q;
  r[0] = f[0]; r[1] = f[1]; r[2] = f[2];
  return dsum;
}

/* permeate: single Powles event per step (geometries.py:1302-1410). */
static float ell_permeate(float r[3], const float st[3], float u, const oracle_args* A,
                          uint32_t* hits, uint32_t* ncross) {
  const float* inv = A->axes;
  const float EPS = A->eps, NUDGE = A->nudge;
  const float sl = norm3(st);
  const float d[3] = {st[0] / sl, st[1] / sl, st[2] / sl};
  const float Aq = ell_q(d, d, inv), Bq = ell_q(r, d, inv), Cq = ell_q(r, r, inv) - 1.0f;
  const float disc_raw = Bq * Bq - Aq * Cq;
  const float sq = sqrtf(fmaxf(disc_raw, 0.0f));
  const int inside = Cq < 0.0f;
  const float t_exit = (-Bq + sq) / Aq, t_entry = (-Bq - sq) / Aq;
  const float th = inside ? t_exit : t_entry;
  const int any = disc_raw > 0.0f && th > EPS && th < sl && sl > 0.0f;
  if (!any) { for (int k = 0; k < 3; ++k) r[k] = r[k] + st[k]; return 0.0f; }
  (*hits)++;
  float h[3], n[3], rf[3];
  for (int k = 0; k < 3; ++k) h[k] = r[k] + th * d[k];
  const float m[3] = {h[0] * inv[0], h[1] * inv[1], h[2] * inv[2]};
  const float mn = norm3(m);
  for (int k = 0; k < 3; ++k) n[k] = m[k] / mn;
  const float remaining = sl - th;
  const float dn = dot3(d, n);
  const float dperp = remaining * fabsf(dn);
  const float p = fminf(1.0f, 2.0f * A->kappa_over_D * dperp);
  if (u < p) { for (int k = 0; k < 3; ++k) r[k] = r[k] + st[k]; (*ncross)++; return 0.0f; }
  for (int k = 0; k < 3; ++k) rf[k] = d[k] - (2.0f * dn) * n[k];
  const float rn = norm3(rf);
  const float go = fmaxf(remaining - NUDGE, 0.0f);
  for (int k = 0; k < 3; ++k) {
    const float nd = inside ? -n[k] : n[k];
    r[k] = (h[k] + NUDGE * nd) + (rf[k] / rn) * go;
  }
  return -2.0f * A->rho_over_D * dperp;
}

/* ---------------------------------------------- PermeableSlab1D (geometries.py:3221-3300) */
static float slab_step(float r[3], const float st[3], float u, const oracle_args* A,
                       uint32_t* hits, uint32_t* ncross) {
  const float L = A->size, xm = 0.5f * A->size;
  const float x = r[0], xn = x + st[0];
  const int crossed = (x - xm) * (xn - xm) < 0.0f;
  const float dperp = crossed ? fabsf(xn - xm) : 0.0f;
  int transmit = 0;
  if (crossed && A->wall == W_PERMEATE) transmit = u < fminf(1.0f, 2.0f * A->kappa_over_D * dperp);
  float x1 = xn, dlog = 0.0f;
  if (crossed) {
    (*hits)++;
    if (transmit) (*ncross)++;
    else {
      x1 = 2.0f * xm - xn;
      if (A->wall != W_REFLECT) dlog = -2.0f * A->rho_over_D * dperp;
    }
  }
  const float L2 = 2.0f * L;
  float mm = fmodf(x1, L2);
  if (mm != 0.0f && ((mm < 0.0f) != (L2 < 0.0f))) mm = mm + L2;
  r[0] = mm > L ? L2 - mm : mm;
  r[1] = r[1] + st[1];
  r[2] = r[2] + st[2];
  return dlog;
}

/* ---------------------------------------------- PermeableShell (geometries.py:3303-3431) */
static void shell_radial(const float v[3], int cylk, float out[3]) {
  out[0] = v[0]; out[1] = v[1]; out[2] = cylk ? 0.0f : v[2];
}

static float shell_first_t(float b, float r2, float dd2, float Rk, float EPS, float sl) {
  const float c = r2 - Rk * Rk;
  const float disc = b * b - dd2 * c;
  const float s = sqrtf(fmaxf(disc, 0.0f));
  float t1 = (-b - s) / dd2, t2 = (-b + s) / dd2;
  if (!(disc > 0.0f && t1 > EPS && t1 < sl)) t1 = 1e30f;
  if (!(disc > 0.0f && t2 > EPS && t2 < sl)) t2 = 1e30f;
  return fminf(t1, t2);
}

static float shell_step(float r[3], const float st[3], float u, const oracle_args* A,
                        uint32_t* hits, uint32_t* ncross) {
  const int cylk = A->kind == K_SHELLC;
  const float Rin = A->size, Rout = A->aux, EPS = A->eps, NUDGE = A->nudge;
  const float sl = norm3(st);
  const float den = fmaxf(sl, EPS);
  const float d[3] = {st[0] / den, st[1] / den, st[2] / den};
  float rr[3], dd[3];
  shell_radial(r, cylk, rr);
  shell_radial(d, cylk, dd);
  const float b = dot3(dd, rr), r2 = dot3(rr, rr), dd2 = fmaxf(dot3(dd, dd), EPS);
  const float t_in = shell_first_t(b, r2, dd2, Rin, EPS, sl);
  const float t_out = shell_first_t(b, r2, dd2, Rout, EPS, sl);
  const int hit_in = t_in < t_out;
  const float th = fminf(t_in, t_out);
  const int any = th < 1e29f;
  float o[3] = {r[0] + st[0], r[1] + st[1], r[2] + st[2]};
  float dlog = 0.0f;
  if (any) {
    (*hits)++;
    float h[3], hr[3], n[3];
    for (int k = 0; k < 3; ++k) h[k] = r[k] + th * d[k];
    shell_radial(h, cylk, hr);
    const float hn = fmaxf(norm3(hr), EPS);
    for (int k = 0; k < 3; ++k) n[k] = hr[k] / hn;
    const float remaining = sl - th;
    const float dn = dot3(d, n);
    const float dperp = remaining * fabsf(dn);
    int transmit = 0;
    if (A->wall == W_PERMEATE && hit_in) transmit = u < fminf(1.0f, 2.0f * A->kappa_over_D * dperp);
    if (transmit) (*ncross)++;
    else {
      const float Rk = hit_in ? Rin : Rout;
      const float side = norm3(rr) < Rk ? -1.0f : 1.0f;
      const float go = fmaxf(remaining - NUDGE, 0.0f);
      for (int k = 0; k < 3; ++k) {
        const float q = d[k] - (2.0f * dn) * n[k];
        o[k] = (h[k] + (side * NUDGE) * n[k]) + q * go;
      }
      if (A->wall == W_PERMEATE && hit_in) dlog = -2.0f * A->rho_over_D * dperp;
    }
  }
  float orad[3];
  shell_radial(o, cylk, orad);
  const float rm = norm3(orad);
  const float over = rm - Rout;
  if (over > 0.0f) {
    const float rs = fmaxf(rm, EPS);
    for (int k = 0; k < 3; ++k) o[k] = o[k] - (2.0f * over) * (orad[k] / rs);
  }
  r[0] = o[0]; r[1] = o[1]; r[2] = o[2];
  return dlog;
}

/* ---------------------------------------------------------- MyelinatedCylinder
   make_myelin_step_fn (physics.py:513-668) in the cylinder frame: the compartment's
   step; crossing flags from the compartment and |r_new_xy|; transmission
   probabilities min(1, 2 kappa/D_leaving |r - R|); inner mirror, then the outer
   mirror on the recomputed norm; the four wrong-side clamps, each on a freshly
   computed norm.  The reference draws separate uniforms for the two membranes; at
   most one membrane is crossed per step, so one uniform u serves both here. */
static float xy_norm(const float r[3]) { return sqrtf(r[0] * r[0] + r[1] * r[1]); }

static void myel_clamp(float r[3], int cond, float Rt) {
  const float n = xy_norm(r);
  const float s = fmaxf(n, 1e-20f);
  if (cond) { r[0] = (r[0] / s) * Rt; r[1] = (r[1] / s) * Rt; }
}

static void myel_step(float r[3], const float dir[3], float u, const oracle_args* A, int* comp,
                      uint32_t* hits, uint32_t* ncross) {
  const int c = *comp;
  const float Rin = A->size, Rout = A->aux, NUDGE = A->nudge;
  const float sl = c == 0 ? A->m_step[0] : (c == 1 ? A->m_step[1] : A->m_step[2]);
  float rn[3] = {r[0] + dir[0] * sl, r[1] + dir[1] * sl, r[2] + dir[2] * sl};
  const float n = xy_norm(rn);
  const int cio = c == 0 && n >= Rin, cii = c == 1 && n < Rin;
  const int coo = c == 1 && n >= Rout, coi = c == 2 && n < Rout;
  const int ci = cio || cii, co = coo || coi;
  const float p_in = fminf(1.0f, (c == 0 ? A->m_k2[0] : A->m_k2[1]) * fabsf(n - Rin));
  const float p_out = fminf(1.0f, (c == 1 ? A->m_k2[2] : A->m_k2[3]) * fabsf(n - Rout));
  const int t_in = ci && (u < p_in), t_out = co && (u < p_out);
  if (ci || co) (*hits)++;
  int nc = c;
  {
    const float s = fmaxf(n, 1e-20f);
    const float hx = rn[0] / s, hy = rn[1] / s;
    const float rr = 2.0f * Rin - n;
    if (ci && !t_in) { rn[0] = hx * rr; rn[1] = hy * rr; }
    if (t_in) nc = cio ? 1 : 0;
  }
  {
    const float n2 = xy_norm(rn);
    const float s = fmaxf(n2, 1e-20f);
    const float hx = rn[0] / s, hy = rn[1] / s;
    const float rr = 2.0f * Rout - n2;
    if (co && !t_out) { rn[0] = hx * rr; rn[1] = hy * rr; }
    if (t_out) nc = coo ? 2 : 1;
  }
  if (t_in || t_out) (*ncross)++;
  myel_clamp(rn, nc == 0 && xy_norm(rn) >= Rin, Rin - NUDGE);
  myel_clamp(rn, nc == 1 && xy_norm(rn) < Rin, Rin + NUDGE);
  myel_clamp(rn, nc == 1 && xy_norm(rn) >= Rout, Rout - NUDGE);
  myel_clamp(rn, nc == 2 && xy_norm(rn) < Rout, Rout + NUDGE);
  r[0] = rn[0]; r[1] = rn[1]; r[2] = rn[2];
  *comp = nc;
}

/* ------------------------------------------------ PackedMyelinatedCylinders
   make_packed_myelin_traj_step_fn (physics.py:726-978, no magnetisation transfer) in
   the reference's compartment encoding: 0 extra, k+1 intra of axon k, n+k+1 myelin
   of axon k (N_max = n here).  Every membrane quantity refers to the walker's own
   axon k_cyl (axon 0 for extra-axonal walkers); the exterior clamp scans every axon.
   One uniform u decides whichever membrane is crossed.  Returns the unit
   (rho/D = 1) boundary log-weight. */
static float wrapL(float q, float L) { return q - L * floorf(q / L + 0.5f); }

static int pm_to3(int cid, int n) { return cid == 0 ? 2 : (cid > n ? 1 : 0); }

static float pm_step_oracle(float r[3], const float dir[3], float u, const oracle_args* A, int* cid,
                            uint32_t* hits, uint32_t* ncross) {
  const int n = A->n_obj;
  const float L = A->cell, NUDGE = A->nudge;
  const int c = *cid;
  const int is_extra = c == 0, is_intra = c >= 1 && c <= n, is_myelin = c > n;
  int k = is_intra ? c - 1 : (is_myelin ? c - (n + 1) : 0);
  if (k < 0) k = 0;
  const float* st = A->pm_step + 3 * k;
  const float sl = is_intra ? st[0] : (is_myelin ? st[1] : st[2]);
  const float step[3] = {dir[0] * sl, dir[1] * sl, dir[2] * sl};
  const float cx = A->centers[2 * k], cy = A->centers[2 * k + 1];
  const float Rin = A->pm_inner[k], Rout = A->radii[k];
  float q[2] = {wrapL((r[0] + step[0]) - cx, L), wrapL((r[1] + step[1]) - cy, L)};
  const float nrm = sqrtf(q[0] * q[0] + q[1] * q[1]);
  const int cio = is_intra && nrm >= Rin, cii = is_myelin && nrm < Rin;
  const int coo = is_myelin && nrm >= Rout, coi = is_extra && nrm < Rout;
  const int ci = cio || cii, co = coo || coi;
  const float dpi = fabsf(nrm - Rin), dpo = fabsf(nrm - Rout);
  const float* k2 = A->pm_k2 + 4 * k;
  const float p_in = fminf(1.0f, (is_intra ? k2[0] : k2[1]) * dpi);
  const float p_out = fminf(1.0f, (is_myelin ? k2[2] : k2[3]) * dpo);
  const int t_in = ci && u < p_in, t_out = co && u < p_out;
  if (ci || co) (*hits)++;
  int nc = c;
  float db = 0.0f;
  {
    const float s = fmaxf(nrm, 1e-20f);
    const float h0 = q[0] / s, h1 = q[1] / s;
    const float rr = 2.0f * Rin - nrm;
    if (ci && !t_in) { q[0] = h0 * rr; q[1] = h1 * rr; }
    if (t_in) nc = cio ? n + k + 1 : k + 1;
    db = db + ((ci && !t_in) ? -2.0f * dpi : 0.0f);
  }
  {
    const float n2 = sqrtf(q[0] * q[0] + q[1] * q[1]);
    const float s = fmaxf(n2, 1e-20f);
    const float h0 = q[0] / s, h1 = q[1] / s;
    const float rr = 2.0f * Rout - n2;
    if (co && !t_out) { q[0] = h0 * rr; q[1] = h1 * rr; }
    if (t_out) nc = coo ? 0 : n + k + 1;
    db = db + ((co && !t_out) ? -2.0f * dpo : 0.0f);
  }
  if (t_in || t_out) (*ncross)++;
  float xy[2] = {wrapL(q[0] + cx, L), wrapL(q[1] + cy, L)};
  /* exterior clamp: argmin over every axon of d^2/R_out^2 inside its outer disc */
  int kcl = 0;
  float best = 0.0f;
  for (int j = 0; j < n; ++j) {
    const float fx = wrapL(xy[0] - A->centers[2 * j], L), fy = wrapL(xy[1] - A->centers[2 * j + 1], L);
    const float d2 = fx * fx + fy * fy;
    const float Ro = A->radii[j];
    const float pen = Ro > 0.0f ? (d2 < Ro * Ro ? d2 / (Ro * Ro + 1e-30f) : 1.0f) : 1.0f;
    if (j == 0 || pen < best) { best = pen; kcl = j; }
  }
  if (is_extra && best < 1.0f) {
    const float Rc = A->radii[kcl];
    const float qx = wrapL(xy[0] - A->centers[2 * kcl], L), qy = wrapL(xy[1] - A->centers[2 * kcl + 1], L);
    const float dcl = sqrtf(qx * qx + qy * qy);
    const float nx = xy[0] - qx, ny = xy[1] - qy;
    const float drefl = 2.0f * Rc - dcl;
    xy[0] = nx + qx * drefl / fmaxf(dcl, NUDGE);
    xy[1] = ny + qy * drefl / fmaxf(dcl, NUDGE);
    db = db + -2.0f * fmaxf(Rc - dcl, 0.0f);
  }
  /* wrong-side clamps of untransmitted intra / myelin walkers */
  const int same = nc == c;
  {
    const float ex = wrapL(xy[0] - cx, L), ey = wrapL(xy[1] - cy, L);
    const float de = sqrtf(ex * ex + ey * ey);
    const float s = fmaxf(de, 1e-20f);
    if (is_intra && same && de >= Rin) { xy[0] = cx + ex / s * (Rin - NUDGE); xy[1] = cy + ey / s * (Rin - NUDGE); }
  }
  {
    const float ex = wrapL(xy[0] - cx, L), ey = wrapL(xy[1] - cy, L);
    const float de = sqrtf(ex * ex + ey * ey);
    const float s = fmaxf(de, 1e-20f);
    if (is_myelin && same && de < Rin) { xy[0] = cx + ex / s * (Rin + NUDGE); xy[1] = cy + ey / s * (Rin + NUDGE); }
  }
  {
    const float ex = wrapL(xy[0] - cx, L), ey = wrapL(xy[1] - cy, L);
    const float de = sqrtf(ex * ex + ey * ey);
    const float s = fmaxf(de, 1e-20f);
    if (is_myelin && same && de >= Rout) { xy[0] = cx + ex / s * (Rout - NUDGE); xy[1] = cy + ey / s * (Rout - NUDGE); }
  }
  r[0] = xy[0];
  r[1] = xy[1];
  r[2] = r[2] + step[2];
  /* intra / myelin walkers re-classified against their own axon */
  const float fx = wrapL(r[0] - cx, L), fy = wrapL(r[1] - cy, L);
  const float d2 = fx * fx + fy * fy;
  if (!is_extra) {
    if (d2 < Rin * Rin) nc = k + 1;
    else if (d2 < Rout * Rout) nc = n + k + 1;
  }
  *cid = nc;
  return db;
}

/* Start placement: segment j of pm_start is intra of axon j (j < n), myelin of axon
   j - n (j < 2n); walkers past the last boundary are extra-axonal, uniform in the
   cell outside every outer disc (geometries.py:3112-3210). */
static int pm_start_oracle(const oracle_args* A, uint32_t w, float r[3]) {
  const int n = A->n_obj;
  const float L = A->cell;
  int cid = 0, seg = -1;
  for (int j = 0; j < 2 * n; ++j)
    if ((int64_t)w >= A->pm_start[j] && (int64_t)w < A->pm_start[j + 1]) seg = j;
  if (seg >= 0) cid = seg < n ? seg + 1 : n + (seg - n) + 1;
  const uint32_t key[2] = {(uint32_t)A->seed, (uint32_t)(A->seed >> 32)};
  r[0] = r[1] = r[2] = 0.0f;
  for (uint32_t att = 0; att < 4096u; ++att) {
    const uint32_t ctr[4] = {w, A->instance, att, 1u};
    uint32_t o[4];
    philox10(ctr, key[0], key[1], o);
    const float u0 = unit_u(o[0]), u1 = unit_u(o[1]);
    if (cid == 0) {
      r[0] = (u0 - 0.5f) * L;
      r[1] = (u1 - 0.5f) * L;
      int out = 1;
      for (int j = 0; j < n; ++j) {
        const float fx = wrapL(r[0] - A->centers[2 * j], L), fy = wrapL(r[1] - A->centers[2 * j + 1], L);
        if (!(fx * fx + fy * fy > A->radii[j] * A->radii[j])) out = 0;
      }
      if (out) return cid;
    } else {
      const int k = cid <= n ? cid - 1 : cid - n - 1;
      const float Rin = A->pm_inner[k], Rout = A->radii[k];
      const float rh = cid <= n ? Rin : Rout;
      const float qx = (2.0f * u0 - 1.0f) * rh, qy = (2.0f * u1 - 1.0f) * rh;
      const float d2 = qx * qx + qy * qy;
      const int ok = cid <= n ? d2 < rh * rh : (d2 >= Rin * Rin && d2 < rh * rh);
      if (ok) {
        r[0] = qx + A->centers[2 * k];
        r[1] = qy + A->centers[2 * k + 1];
        return cid;
      }
    }
  }
  return cid;
}

/* classify_position of MyelinatedCylinder (geometries.py:1035-1052). */
static int myel_classify(const float r[3], const oracle_args* A) {
  const float rr = r[0] * r[0] + r[1] * r[1];
  return rr < A->size * A->size ? 0 : (rr < A->aux * A->aux ? 1 : 2);
}

/* Producer compartment occupancy rule (dmipy_sim/core.py:1166-1185): keyed on the
   geometry's `radius`; Cylinder by the axial-plane norm, Sphere / slab (radius L/2) /
   shells (radius R_in) by the 3-D norm; no radius (Ellipsoid, others): 0. */
static float occupancy_outside(const float r[3], const oracle_args* A) {
  const float r2 = sqrtf(r[0] * r[0] + r[1] * r[1]);
  const float r3 = sqrtf((r[0] * r[0] + r[1] * r[1]) + r[2] * r[2]);
  switch (A->kind) {
    case K_CYL: return r2 < A->size ? 0.0f : 1.0f;
    case K_SPH: case K_SHELL: case K_SHELLC: return r3 < A->size ? 0.0f : 1.0f;
    case K_SLAB: return r3 < 0.5f * A->size ? 0.0f : 1.0f;
    default: return 0.0f;
  }
}

static int classify(const float r[3], const oracle_args* A) {
  switch (A->kind) {
    case K_CYL: return (r[0] * r[0] + r[1] * r[1]) < A->size * A->size ? 0 : 1;
    case K_SPH: return dot3(r, r) < A->size * A->size ? 0 : 1;
    case K_ELL: return ell_q(r, r, A->axes) < 1.0f ? 0 : 1;
    case K_SLAB: return r[0] < 0.5f * A->size ? 0 : 1;
    case K_SHELL:
    case K_SHELLC: {
      float rr[3];
      shell_radial(r, A->kind == K_SHELLC, rr);
      return norm3(rr) < A->size ? 0 : 1;
    }
    case K_PCYL:
    case K_PSPH: {
      const int dim = A->kind == K_PCYL ? 2 : 3;
      int id = 0;
      for (int k = 0; k < A->n_obj; ++k) {
        float c[3], q[3] = {0, 0, 0};
        obj_center(A, k, dim, c);
        for (int d = 0; d < dim; ++d) q[d] = mi(r[d] - c[d], A->cell);
        if (dotd(q, q, dim) < A->radii[k] * A->radii[k] && k + 1 > id) id = k + 1;
      }
      return id;
    }
    default: return 0;
  }
}

/* Start position sampler: rejection from Philox draws with tag 1. */
static void sample_start(const oracle_args* A, uint32_t w, float r[3]) {
  r[0] = r[1] = r[2] = 0.0f;
  if (A->kind == K_FREE) return;
  const uint32_t key[2] = {(uint32_t)A->seed, (uint32_t)(A->seed >> 32)};
  for (uint32_t att = 0; att < 4096u; ++att) {
    const uint32_t ctr[4] = {w, A->instance, att, 1u};
    uint32_t o[4];
    philox10(ctr, key[0], key[1], o);
    const float u0 = unit_u(o[0]), u1 = unit_u(o[1]), u2 = unit_u(o[2]);
    if (A->kind == K_BOX) { r[0] = u0 * A->size; return; }
    if (A->kind == K_CYL) {
      r[0] = (2.0f * u0 - 1.0f) * A->size; r[1] = (2.0f * u1 - 1.0f) * A->size;
      if (r[0] * r[0] + r[1] * r[1] < A->size * A->size) return;
    } else if (A->kind == K_SPH || A->kind == K_SHELL) {
      r[0] = (2.0f * u0 - 1.0f) * A->size; r[1] = (2.0f * u1 - 1.0f) * A->size;
      r[2] = (2.0f * u2 - 1.0f) * A->size;
      if (dot3(r, r) < A->size * A->size) return;
    } else if (A->kind == K_SHELLC) {
      r[0] = (2.0f * u0 - 1.0f) * A->size; r[1] = (2.0f * u1 - 1.0f) * A->size;
      r[2] = u2 * A->size;
      if (r[0] * r[0] + r[1] * r[1] < A->size * A->size) return;
    } else if (A->kind == K_SLAB) {
      r[0] = u0 * (0.5f * A->size); r[1] = 0.0f; r[2] = 0.0f;
      return;
    } else if (A->kind == K_MYEL) {
      /* uniform in the start compartment's annulus, the extra-axonal one bounded
         at 2 R_out (geometries.py:921-1008); compartment by walker index */
      const int c0 = (int64_t)w < A->m_n0 ? 0 : ((int64_t)w < A->m_n01 ? 1 : 2);
      const float rh = c0 == 0 ? A->size : (c0 == 1 ? A->aux : 2.0f * A->aux);
      r[0] = (2.0f * u0 - 1.0f) * rh; r[1] = (2.0f * u1 - 1.0f) * rh; r[2] = 0.0f;
      if (myel_classify(r, A) == c0 && r[0] * r[0] + r[1] * r[1] < rh * rh) return;
    } else if (A->kind == K_ELL) {
      const float b[3] = {2.0f * u0 - 1.0f, 2.0f * u1 - 1.0f, 2.0f * u2 - 1.0f};
      for (int k = 0; k < 3; ++k) r[k] = b[k] / sqrtf(A->axes[k]);
      if (dot3(b, b) < 1.0f && classify(r, A) == 0) return;
    } else {
      r[0] = (u0 - 0.5f) * A->cell; r[1] = (u1 - 0.5f) * A->cell;
      r[2] = A->kind == K_PSPH ? (u2 - 0.5f) * A->cell : 0.0f;
      if (classify(r, A) == 0) return;
    }
  }
}

/* One sub-step: direction, geometry, returns dlog (before chi gating). */
static float substep(float r[3], const float dir[3], float u, const oracle_args* A,
                     uint32_t* hits, uint32_t* ncross) {
  const float st[3] = {dir[0] * A->step_len, dir[1] * A->step_len, dir[2] * A->step_len};
  switch (A->kind) {
    case K_FREE:
      r[0] = r[0] + st[0]; r[1] = r[1] + st[1]; r[2] = r[2] + st[2];
      return 0.0f;
    case K_BOX: return box_step(r, st, A->size, A->rho_over_D, A->wall == W_SURFACE, hits);
    case K_CYL:
      if (A->wall == W_PERMEATE) return cyl_permeate(r, st, u, A, hits, ncross);
      return -2.0f * A->rho_over_D * cyl_reflect(r, st, A, hits);
    case K_SPH:
      if (A->wall == W_PERMEATE) return sph_permeate(r, st, u, A, hits, ncross);
      return -2.0f * A->rho_over_D * sph_reflect(r, st, A, hits);
    case K_ELL:
      if (A->wall == W_PERMEATE) return ell_permeate(r, st, u, A, hits, ncross);
      return A->wall == W_SURFACE ? -2.0f * A->rho_over_D * ell_reflect(r, st, A, hits)
                                  : (ell_reflect(r, st, A, hits), 0.0f);
    case K_SLAB: return slab_step(r, st, u, A, hits, ncross);
    case K_SHELL:
    case K_SHELLC: return shell_step(r, st, u, A, hits, ncross);
    default: {
      const int dim = A->kind == K_PCYL ? 2 : 3;
      if (A->wall == W_PERMEATE) return packed_permeate(r, st, u, A, dim, hits, ncross);
      return -2.0f * A->rho_over_D * packed_reflect(r, st, A, dim, hits);
    }
  }
}

int oracle_walk(const oracle_args* A, oracle_out* O) {
  const int64_t N = A->n_walkers;
  const int64_t w0 = A->walker_begin, w1 = A->walker_end > 0 ? A->walker_end : N;
  const int M = A->n_meas, T = A->n_t, NS = A->n_sub;
  const int64_t S = (int64_t)T * NS;
  const int rows = A->n_echo > 0 ? A->n_echo : 1;
  const int packed = A->kind == K_PCYL || A->kind == K_PSPH;
  const int dim = A->kind == K_PSPH ? 3 : 2;
  const float L = A->cell;
  const uint32_t key[2] = {(uint32_t)A->seed, (uint32_t)(A->seed >> 32)};
  double* acc = (double*)calloc((size_t)rows * M * 2, sizeof(double));
  double* accx = (double*)calloc((size_t)rows * M * 2, sizeof(double));
  const int want_x = O->signal_x != NULL;
  uint64_t hits_total = 0;
  double wsum = 0.0;      /* sum of spin weights (_ens denominator, core.py:426-429) */
  if (!acc || !accx) return 1;
#ifdef _OPENMP
  if (A->threads > 0) omp_set_num_threads(A->threads);
#endif
#pragma omp parallel
  {
    double* loc = (double*)calloc((size_t)rows * M * 2, sizeof(double));
    float* phi = (float*)malloc(sizeof(float) * (M > 0 ? M : 1));
    double* phi64 = (double*)malloc(sizeof(double) * (M > 0 ? M : 1));
    double* phix = (double*)malloc(sizeof(double) * (M > 0 ? M : 1));
    double* locx = (double*)calloc((size_t)rows * M * 2, sizeof(double));
    uint64_t hits_loc = 0;
    double wsum_loc = 0.0;
#pragma omp for schedule(dynamic, 1)
    for (int64_t w = w0; w < w1; ++w) {
      float r[3];
      int cell[3] = {0, 0, 0};
      const int pmyel = A->kind == K_PMYEL;
      int pcid = 0;                                   /* packed myelin: compartment id */
      const uint32_t wg = (uint32_t)(w + A->first_walker);   /* global walker index */
      if (pmyel) pcid = pm_start_oracle(A, wg, r);
      else if (A->r0) { r[0] = A->r0[w * 3]; r[1] = A->r0[w * 3 + 1]; r[2] = A->r0[w * 3 + 2]; }
      else sample_start(A, wg, r);
      const int pc0 = pm_to3(pcid, A->n_obj);
      float uw[3] = {r[0], r[1], r[2]};               /* packed myelin: unwrapped position */
      if (packed && A->cell_form) {
        for (int d = 0; d < dim; ++d) {
          const float k = floorf(r[d] / L + 0.5f);
          if (k != 0.0f) { r[d] = r[d] - L * k; cell[d] += (int)k; }
        }
      }
      const int myel = A->kind == K_MYEL;
      int mstate = myel ? myel_classify(r, A) : 0;   /* carried compartment */
      const int c0 = mstate;
      int comp = myel ? mstate : (pmyel ? pc0 : classify(r, A));
      if (O->comp_origin) O->comp_origin[w] = comp;
      for (int m = 0; m < M; ++m) { phi[m] = 0.0f; phi64[m] = 0.0; phix[m] = 0.0; }
      float logw = 0.0f;
      uint32_t hits = 0, ncross = 0;
      uint32_t rnd[4] = {0, 0, 0, 0};
      int e_next = 0;
      for (int t = 0; t < T; ++t) {
        const float chi = A->chi ? A->chi[t] : 1.0f;
        float dlog_sample = 0.0f, occ = 0.0f;
        for (int s = 0; s < NS; ++s) {
          const int64_t gs = (int64_t)t * NS + s;
          float dir[3], u = 0.0f;
          if (A->inj_n) {
            const float* nv = A->inj_n + ((size_t)w * S + gs) * 3;
            const float nn = sqrtf((nv[0] * nv[0] + nv[1] * nv[1]) + nv[2] * nv[2]);
            dir[0] = nv[0] / nn; dir[1] = nv[1] / nn; dir[2] = nv[2] / nn;
            if (A->inj_u) u = A->inj_u[(size_t)w * S + gs];
          } else {
            if ((gs & 1) == 0 || gs == 0) {
              const uint32_t ctr[4] = {wg, A->instance, (uint32_t)(gs >> 1), 0u};
              philox10(ctr, key[0], key[1], rnd);
            }
            oracle_direction((gs & 1) ? rnd[2] : rnd[0], (gs & 1) ? rnd[3] : rnd[1], dir);
            if (A->wall == W_PERMEATE || myel || pmyel) {
              const uint32_t ctr[4] = {wg, A->instance, (uint32_t)gs, 2u};
              uint32_t o[4];
              philox10(ctr, key[0], key[1], o);
              u = unit_u(o[0]);
            }
          }
          float dlog = 0.0f;
          if (pmyel) {
            /* make_packed_myelin_step_fn (physics.py:1033-1060): the walk, the unwrapped
               displacement, then T2 of the new class, T1, and the walls' relaxivity */
            const float r_old[3] = {r[0], r[1], r[2]};
            const float db = pm_step_oracle(r, dir, u, A, &pcid, &hits, &ncross);
            float dr[3] = {r[0] - r_old[0], r[1] - r_old[1], r[2] - r_old[2]};
            for (int d = 0; d < 2; ++d) dr[d] = dr[d] - L * rintf(dr[d] / L);
            for (int d = 0; d < 3; ++d) uw[d] = uw[d] + dr[d];
            if (A->has_t2) dlog = dlog - A->pm_q2[pm_to3(pcid, A->n_obj)] * chi;
            if (A->has_t1) dlog = dlog - A->dt_sub * A->inv_T1 * (1.0f - chi);
            if (A->pm_rhoD > 0.0f) dlog = dlog + A->pm_rhoD * db * chi;
          } else {
            if (myel) myel_step(r, dir, u, A, &mstate, &hits, &ncross);
            else dlog = substep(r, dir, u, A, &hits, &ncross);
            if (A->wall != W_REFLECT) dlog = dlog * chi;
            if (A->has_t2) {
              /* myelin: dt / T2 of the compartment after the step (physics.py:657-659) */
              const float q2 = myel ? A->m_q2[mstate] : A->dt_sub * A->inv_T2;
              dlog = dlog - q2 * chi;
            }
            if (A->has_t1) dlog = dlog - (A->dt_sub * A->inv_T1) * (1.0f - chi);
          }
          logw = logw + dlog;
          dlog_sample = dlog_sample + dlog;
          if (O->occ_full) occ = occ + occupancy_outside(r, A);
          /* phase of every measurement, right-endpoint rule */
          float ru[3] = {r[0], r[1], r[2]};
          if (pmyel) { ru[0] = uw[0]; ru[1] = uw[1]; ru[2] = uw[2]; }
          if (packed && A->cell_form)
            for (int d = 0; d < 3; ++d) ru[d] = r[d] + L * (float)cell[d];
          for (int m = 0; m < M; ++m) {
            const float* g = A->G + ((size_t)m * T + t) * 3;
            const float gr = (g[0] * ru[0] + g[1] * ru[1]) + g[2] * ru[2];
            if (A->phase64) phi64[m] += (double)A->gamma_dt_sub * (double)gr;
            else phi[m] = phi[m] + A->gamma_dt_sub * gr;
          }
          if (want_x) {
            double rd[3];
            for (int d = 0; d < 3; ++d) rd[d] = (double)ru[d];
            if (packed && A->cell_form)
              for (int d = 0; d < 3; ++d) rd[d] = (double)r[d] + (double)L * (double)cell[d];
            for (int m = 0; m < M; ++m) {
              const float* g = A->G + ((size_t)m * T + t) * 3;
              const double gr = ((double)g[0] * rd[0] + (double)g[1] * rd[1]) + (double)g[2] * rd[2];
              phix[m] += (double)A->gamma_dt_sub * gr;
            }
          }
        }
        if (packed && A->cell_form) {
          for (int d = 0; d < dim; ++d) {
            const float k = floorf(r[d] / L + 0.5f);
            if (k != 0.0f) { r[d] = r[d] - L * k; cell[d] += (int)k; }
          }
        }
        if (O->pos_full)
          for (int d = 0; d < 3; ++d)
            O->pos_full[((size_t)w * T + t) * 3 + d] = (packed && A->cell_form) ? r[d] + L * (float)cell[d] : r[d];
        if (O->dlog_full) O->dlog_full[(size_t)w * T + t] = dlog_sample;
        if (O->occ_full) O->occ_full[(size_t)w * T + t] = occ / (float)NS;
        if (O->comp_full || (O->comp_final && t == T - 1))
          comp = myel ? mstate : (pmyel ? pm_to3(pcid, A->n_obj) : classify(r, A));
        if (O->comp_full) O->comp_full[(size_t)w * T + t] = comp;
        if (e_next < A->n_echo && A->echo_idx[e_next] == t) {
          const double wg = exp((double)logw) * (myel ? (double)A->m_wf[c0] : 1.0) *
                            ((pmyel && pc0 == 1) ? (double)A->pm_wmy : 1.0);
          for (int m = 0; m < M; ++m) {
            const double ph = A->phase64 ? phi64[m] : (double)phi[m];
            loc[((size_t)e_next * M + m) * 2] += wg * cos(ph);
            loc[((size_t)e_next * M + m) * 2 + 1] += wg * sin(ph);
            if (want_x) {
              locx[((size_t)e_next * M + m) * 2] += wg * cos(phix[m]);
              locx[((size_t)e_next * M + m) * 2 + 1] += wg * sin(phix[m]);
            }
          }
          ++e_next;
        }
      }
      if (A->n_echo == 0) {
        const double wg = exp((double)logw) * (myel ? (double)A->m_wf[c0] : 1.0) *
                          ((pmyel && pc0 == 1) ? (double)A->pm_wmy : 1.0);
        for (int m = 0; m < M; ++m) {
          const double ph = A->phase64 ? phi64[m] : (double)phi[m];
          loc[m * 2] += wg * cos(ph);
          loc[m * 2 + 1] += wg * sin(ph);
          if (want_x) {
            locx[m * 2] += wg * cos(phix[m]);
            locx[m * 2 + 1] += wg * sin(phix[m]);
          }
        }
      }
      if (O->final_pos)
        for (int d = 0; d < 3; ++d)
          O->final_pos[w * 3 + d] = (packed && A->cell_form) ? r[d] + L * (float)cell[d] : r[d];
      if (O->comp_final) O->comp_final[w] = comp;
      if (O->n_cross) O->n_cross[w] = ncross;
      if (O->log_w) O->log_w[w] = logw;
      if (O->phi)
        for (int m = 0; m < M; ++m) O->phi[w * M + m] = A->phase64 ? (float)phi64[m] : phi[m];
      if (O->phi_x)
        for (int m = 0; m < M; ++m) O->phi_x[w * M + m] = phix[m];
      hits_loc += hits;
      wsum_loc += A->kind == K_MYEL ? (double)A->m_wf[c0]
                                    : ((pmyel && pc0 == 1) ? (double)A->pm_wmy : 1.0);
    }
#pragma omp critical
    {
      for (int i = 0; i < rows * M * 2; ++i) acc[i] += loc[i];
      for (int i = 0; i < rows * M * 2; ++i) accx[i] += locx[i];
      hits_total += hits_loc;
      wsum += wsum_loc;
    }
    free(loc);
    free(phi);
    free(phi64);
    free(phix);
    free(locx);
  }
  const double inv = 1.0 / wsum;
  for (int i = 0; i < rows * M; ++i) {
    O->signal[i] = acc[2 * i] * inv;
    if (O->signal_im) O->signal_im[i] = acc[2 * i + 1] * inv;
    if (want_x) {
      O->signal_x[i] = accx[2 * i] * inv;
      if (O->signal_x_im) O->signal_x_im[i] = accx[2 * i + 1] * inv;
    }
  }
  O->n_hits = hits_total;
  free(acc);
  free(accx);
  return 0;
}
