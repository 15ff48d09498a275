"""Pins the oracle's permeable walls (Powles transmission) against float64 geometry.

For every permeable wall -- Cylinder, Sphere, Ellipsoid, PackedCylinders (also
through a periodic image), PackedSpheres -- a walker starts on either side of the
wall and one step of known length crosses it at a known point with incidence
angle 0 or 36.87 degrees (cos 0.8).  The expected event is solved in float64 from
the walker's float32 start and step, independently of the oracle's arithmetic:

  hit point      first root of the ray with the wall (exit from inside, entry
                 from outside), over every object and periodic image
  d_perp         how far the straight end point lies beyond the wall's tangent
                 plane at the hit point: remaining length x |cos(incidence)|
                 (the axial plane's for cylinders)
  transmission   u < min(1, 2 kappa / D d_perp)   (geometries.py:320-424,
                 656-801, 1302-1410, 2018-2183, 2633-2756)
  transmit       the straight end point, a compartment change, no weight
  mirror         the tangent-plane mirror of the remaining path, nudged to the
                 walker's side; log-weight -2 rho / D d_perp; compartment kept

Each case runs with u a relative 1e-3 below and above the float64 probability, so
a missing factor 2, a missing incidence cosine, the wrong side's root, a wrong
nudge sign or a swapped compartment fails.  The PackedSpheres reflection test
mirrors the PackedCylinders one in test_oracle.py: a float64 ray/sphere solution
over the 27 periodic images of every sphere.
"""
import numpy as np
import pytest

import dmipy_sim_amd as ds
from substrates import D, packed_sph

f32 = np.float32
ALPHAS = [0.0, float(np.arccos(0.8))]   # incidence angles
REL = 1e-3                               # u = p (1 -/+ REL)


def _wave(dt=1e-4):
    return ds.Waveform(G=np.zeros((1, 1, 3), np.float32), dt=dt, echo_idx=0)


def _run(oracle, geom, r0, u3, step, kD, rD, u):
    wf = _wave()
    plan = oracle.oracle_plan(geom, D, wf.dt)
    plan.update(n_sub=1, dt_sub=float(f32(wf.dt)), step_len=float(step), kappa_over_D=float(kD),
                rho_over_D=float(rD))
    res = oracle.run_oracle(geom, D, wf, 1, r0=np.asarray(r0, f32).reshape(1, 3), plan=plan,
                            inject_normals=np.asarray(u3, f32).reshape(1, 1, 3),
                            inject_uniforms=np.asarray([[u]], f32))
    return res, plan


def _perp(n, rng):
    t = rng.standard_normal(len(n))
    t -= t.dot(n) * n
    return t / np.linalg.norm(t)


# --------------------------------------------------------------- wall models
class Round:
    """Circle (dim 2, axial plane of a z cylinder) or sphere (dim 3) of radius R
    around centre c (periodic images of a packing are separate Round walls)."""

    def __init__(self, c, R, dim):
        self.c, self.R, self.dim = np.asarray(c, float), float(R), dim

    def point_normal(self, rng):
        n = rng.standard_normal(self.dim)
        n /= np.linalg.norm(n)
        return self.c + self.R * n, n

    def roots(self, p, d):
        q = p - self.c
        b, c0 = d @ q, q @ q - self.R ** 2
        disc = b * b - c0
        if disc <= 0:
            return []
        s = np.sqrt(disc)
        return [(-b - s, c0 < 0), (-b + s, c0 < 0)]

    def normal(self, h):
        return (h - self.c) / self.R

    def inside(self, p):
        return (p - self.c) @ (p - self.c) < self.R ** 2


class Ell:
    """Ellipsoid with semi-axes a (centred at the origin)."""
    dim = 3

    def __init__(self, a):
        self.a2 = np.asarray(a, float) ** 2
        self.a = np.asarray(a, float)

    def point_normal(self, rng):
        v = rng.standard_normal(3)
        v /= np.linalg.norm(v)
        h = v * self.a
        return h, self.normal(h)

    def normal(self, h):
        m = h / self.a2
        return m / np.linalg.norm(m)

    def roots(self, p, d):
        A, B, C = (d * d / self.a2).sum(), (p * d / self.a2).sum(), (p * p / self.a2).sum() - 1
        disc = B * B - A * C
        if disc <= 0:
            return []
        s = np.sqrt(disc)
        return [((-B - s) / A, C < 0), ((-B + s) / A, C < 0)]

    def inside(self, p):
        return (p * p / self.a2).sum() < 1


def _first_event(walls, p, d, sl, eps):
    """First valid root over the walls: (t, wall, inside) as the reference picks it
    (exit root from inside, entry root from outside, in (eps, sl))."""
    best = None
    for w in walls:
        rts = w.roots(p, d)
        if not rts:
            continue
        inside = rts[0][1]
        t = rts[1][0] if inside else rts[0][0]
        if eps < t < sl and (best is None or t < best[0]):
            best = (t, w, inside)
    return best


# ------------------------------------------------------------------ the cases
def _packed_cyl():
    L = 16e-6
    radii = np.array([3e-6, 2e-6])
    centers = np.array([[1e-6, 2e-6], [0.5 * L - 1.2e-6, -4e-6]])
    return ds.PackedCylinders(radii, centers, L, permeability=1e-5, surface_relaxivity_t2=1e-5)


def _packed_sph():
    L = 14e-6
    radii = np.array([3e-6, 2.2e-6])
    centers = np.array([[1e-6, 1e-6, -1e-6], [0.5 * L - 1.0e-6, -3.5e-6, 3e-6]])
    return ds.PackedSpheres(radii, centers, L, permeability=1e-5, surface_relaxivity_t2=1e-5)


def _images(geom, dim):
    L = geom._L_float
    out = []
    for k, (c, R) in enumerate(zip(geom._centers_np.astype(float), geom._radii_np.astype(float))):
        for off in np.ndindex(*(3,) * dim):
            w = Round(c + L * (np.array(off) - 1), R, dim)
            w.k = k
            out.append(w)
    return out


def _case(name, seed):
    """-> geom, walls (float64), the wall to hit, comp labels (inside, outside), dim."""
    rng = np.random.default_rng(seed)
    if name == "cylinder":
        g = ds.Cylinder(3e-6, (0, 0, 1), permeability=1e-5, surface_relaxivity_t2=1e-5)
        w = Round([0, 0], 3e-6, 2)
        return g, [w], w, (0, 1), rng
    if name == "sphere":
        g = ds.Sphere(3e-6, permeability=1e-5, surface_relaxivity_t2=1e-5)
        w = Round([0, 0, 0], 3e-6, 3)
        return g, [w], w, (0, 1), rng
    if name == "ellipsoid":
        g = ds.Ellipsoid((3e-6, 2e-6, 4e-6), permeability=1e-5, surface_relaxivity_t2=1e-5)
        w = Ell((3e-6, 2e-6, 4e-6))
        return g, [w], w, (0, 1), rng
    if name in ("packed_cyl", "packed_cyl_image"):
        g = _packed_cyl()
        walls = _images(g, 2)
        k = 0 if name == "packed_cyl" else 1
        # object 1 sits by the +x face: hit its image across the -x face
        shift = np.zeros(2) if k == 0 else np.array([-g._L_float, 0.0])
        w = next(x for x in walls if x.k == k and np.allclose(x.c, g._centers_np[k] + shift))
        return g, walls, w, (k + 1, 0), rng
    if name in ("packed_sph", "packed_sph_image"):
        g = _packed_sph()
        walls = _images(g, 3)
        k = 0 if name == "packed_sph" else 1
        shift = np.zeros(3) if k == 0 else np.array([-g._L_float, 0.0, 0.0])
        w = next(x for x in walls if x.k == k and np.allclose(x.c, g._centers_np[k] + shift))
        return g, walls, w, (k + 1, 0), rng
    raise ValueError(name)


NAMES = ["cylinder", "sphere", "ellipsoid", "packed_cyl", "packed_cyl_image", "packed_sph",
         "packed_sph_image"]


@pytest.mark.parametrize("alpha", ALPHAS, ids=["normal", "oblique"])
@pytest.mark.parametrize("side", ["inside", "outside"])
@pytest.mark.parametrize("name", NAMES)
def test_powles_event_matches_float64_geometry(oracle, name, side, alpha):
    seed = 10 * NAMES.index(name) + (side == "inside") + 2 * ALPHAS.index(alpha)
    geom, walls, wall, (c_in, c_out), rng = _case(name, seed)
    dim = wall.dim
    step = 0.15 * 2e-6
    checked = 0
    for trial in range(40):
        h, n = wall.point_normal(rng)
        # direction in the wall's dimensions: cos(incidence) along +n (exit) or -n (entry)
        t_vec = _perp(n, rng)
        dw = np.cos(alpha) * n + np.sin(alpha) * t_vec
        if side == "outside":
            dw = -np.cos(alpha) * n + np.sin(alpha) * t_vec
        beta = 0.45 if dim == 2 else 0.0          # cylinders: a step with an axial part
        d3 = np.append(dw * np.cos(beta), np.sin(beta)) if dim == 2 else dw
        sl_w = step * np.cos(beta)                # step length in the wall's dimensions
        t0 = 0.5 * sl_w
        start = h - t0 * dw
        r0 = np.zeros(3)
        r0[:dim] = start
        if dim == 2:
            r0[2] = 1e-6
        r0 = r0.astype(f32).astype(float)
        # the float32 step the oracle takes: normal / |normal| x step_len
        u3 = d3.astype(f32).astype(float)
        s32 = float(f32(step))
        st = (u3 / np.linalg.norm(u3)) * s32
        p0 = r0[:dim]
        sw = st[:dim]
        slw = np.linalg.norm(sw)
        dh = sw / slw
        plan_eps = 1e-7 * 2e-6
        ev = _first_event(walls, p0, dh, slw, plan_eps)
        assert ev is not None and ev[1] is wall and ev[2] == (side == "inside")
        th, _, inside = ev
        hit = p0 + th * dh
        nn = wall.normal(hit)
        rem = slw - th
        dperp = rem * abs(dh @ nn)
        assert dperp > 0.1 * rem
        kD = float(f32(0.25 / dperp))             # p = 2 kD d_perp ~ 0.5
        rD = 400.0
        p = min(1.0, 2.0 * float(f32(kD)) * dperp)
        # transmission: the straight end point, no weight, the other compartment
        res, plan = _run(oracle, geom, r0, u3, s32, kD, rD, p * (1 - REL))
        assert res["n_cross"][0] == 1
        assert res["comp_origin"][0] == (c_in if inside else c_out)
        assert res["comp_final"][0] == (c_out if inside else c_in)
        np.testing.assert_allclose(res["final_pos"][0].astype(float), r0 + st, rtol=0, atol=2e-12)
        assert res["log_w"][0] == 0.0
        # mirror: tangent-plane reflection of the remaining path, nudged to the walker's side
        res, plan = _run(oracle, geom, r0, u3, s32, kD, rD, p * (1 + REL))
        nudge = plan["nudge"]
        rf = dh - 2 * (dh @ nn) * nn
        rf /= np.linalg.norm(rf)
        side_n = -nn if inside else nn
        expect_w = hit + nudge * side_n + rf * max(rem - nudge, 0.0)
        if isinstance(wall, Round) and (np.linalg.norm(expect_w - wall.c) >= wall.R) == inside:
            continue   # the wrong-side clamp would act: not a plain mirror
        expect = r0 + st
        expect[:dim] = expect_w
        assert res["n_cross"][0] == 0
        assert res["comp_final"][0] == res["comp_origin"][0]
        np.testing.assert_allclose(res["final_pos"][0].astype(float), expect, rtol=0, atol=2e-11)
        np.testing.assert_allclose(float(res["log_w"][0]), -2.0 * float(f32(rD)) * dperp, rtol=2e-4)
        checked += 1
    assert checked >= 30


def _ray_sphere_hits(p, d, centers, radii, L, tmax):
    """float64 brute force over the 27 periodic images: sorted (t, k, disc) entry
    roots along a ray segment of length tmax."""
    hits = []
    for k, (c, R) in enumerate(zip(centers, radii)):
        for off in np.ndindex(3, 3, 3):
            q = p - (c + L * (np.array(off) - 1))
            b = np.dot(d, q)
            disc = b * b - (np.dot(q, q) - R * R)
            if disc <= 0:
                continue
            t = -b - np.sqrt(disc)
            if 0 < t < tmax:
                hits.append((t, k, disc))
    return sorted(hits)


def test_packed_spheres_reflection_matches_float64_geometry(oracle):
    """Reflection off a sphere packing: the float32 oracle matches a float64
    ray/sphere solution of the same step (hit point, specular direction, nudge)
    to 1e-10 m, and its returned d_perp (the surface log-weight) to 1e-4."""
    geom = packed_sph()
    wf = _wave()
    plan = oracle.oracle_plan(geom, D, wf.dt)
    step = 1.5e-6
    rho = 2e-5
    geom_r = packed_sph(rho=rho)
    rng = np.random.default_rng(1)
    L = geom._L_float
    C, Rs = geom._centers_np.astype(float), geom._radii_np.astype(float)
    checked = 0
    for trial in range(600):
        r0 = geom.init_positions(1, trial)[0].astype(np.float64)
        nrm = rng.standard_normal(3)
        u = (nrm.astype(f32) / np.linalg.norm(nrm.astype(f32))).astype(float)
        s = float(f32(step))
        hits = _ray_sphere_hits(r0, u, C, Rs, L, s)
        if not hits:
            continue
        t, k, disc = hits[0]
        if disc < 1e-16 or s - t < 1e-9 or (len(hits) > 1 and hits[1][0] - t < 1e-9):
            continue   # float32/float64 near-ties
        hit = r0 + t * u
        q = hit - C[k]
        q -= L * np.round(q / L)
        n = q / np.linalg.norm(q)
        refl = u - 2 * np.dot(u, n) * n
        nudge = plan["nudge"]
        expect = hit + nudge * n + refl * max(s - t - nudge, 0.0)
        qe = expect - C
        qe -= L * np.round(qe / L)
        if np.any(np.linalg.norm(qe, axis=1) < Rs * (1 + 1e-6)):
            continue   # the safety clamp would act: not a single-reflection case
        wp = dict(plan, n_sub=1, dt_sub=float(f32(wf.dt)), step_len=s)
        res = oracle.run_oracle(geom, D, wf, 1, r0=r0.astype(f32).reshape(1, 3), plan=wp,
                                inject_normals=nrm.astype(f32).reshape(1, 1, 3))
        np.testing.assert_allclose(res["final_pos"][0], expect, atol=1e-10)
        assert res["comp_final"][0] == 0
        # the surface weight: -2 rho/D d_perp, d_perp = (s - t) cos(incidence)
        pr = oracle.oracle_plan(geom_r, D, wf.dt)
        wr = dict(pr, n_sub=1, dt_sub=float(f32(wf.dt)), step_len=s)
        rr = oracle.run_oracle(geom_r, D, wf, 1, r0=r0.astype(f32).reshape(1, 3), plan=wr,
                               inject_normals=nrm.astype(f32).reshape(1, 1, 3))
        dperp = (s - t) * abs(np.dot(u, n))
        np.testing.assert_allclose(float(rr["log_w"][0]), -2 * pr["rho_over_D"] * dperp, rtol=1e-4)
        checked += 1
    assert checked > 20
