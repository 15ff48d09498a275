"""Pins the CPU oracle (walk_oracle.c) before it is trusted as the GPU's checker.

Every check here is independent of the oracle's own formulas: published
known-answer vectors, textbook geometry (normal incidence, slab folding, ray/circle
roots solved in float64), invariants (containment, weights only on reflection)
and closed-form Monte-Carlo limits (exp(-bD), exp(-t/T2), Brownstein-Tarr).
"""
import numpy as np
import pytest

import dmipy_sim_amd as ds
from substrates import D, packed_cyl, packed_sph

GAMMA = ds.GAMMA


# --------------------------------------------------------------- Philox KAT
# Known-answer vectors of Philox4x32-10 published with the Random123 library
# (Salmon et al., SC'11): (counter, key) -> output.
KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
]


@pytest.mark.parametrize("ctr,key,expect", KAT)
def test_philox_known_answers(oracle, ctr, key, expect):
    assert [int(v) for v in oracle.philox4x32_10(ctr, key)] == list(expect)


def _wave(n_t=1, n_meas=1, dt=1e-4):
    return ds.Waveform(G=np.zeros((n_meas, n_t, 3), np.float32), dt=dt, echo_idx=n_t - 1)


def _one_step(oracle, geom, r0, normal, u=None, T2=None, wall_plan=None, per_sample=False):
    """Single walker, single sub-step probe of the oracle's step."""
    wf = _wave()
    plan = oracle.oracle_plan(geom, D, wf.dt, T2)
    plan["n_sub"] = 1
    plan["dt_sub"] = float(np.float32(wf.dt))
    if wall_plan:
        plan.update(wall_plan)
    kw = {"inject_normals": np.asarray(normal, np.float32).reshape(1, 1, 3)}
    if u is not None:
        kw["inject_uniforms"] = np.asarray([[u]], np.float32)
    res = oracle.run_oracle(geom, D, wf, 1, r0=np.asarray(r0, np.float32).reshape(1, 3),
                            plan=plan, per_sample=per_sample, **kw)
    return res, plan


def test_direction_sampler_matches_exact_trig(oracle):
    """Philox-driven directions: unit norm and equal (to 1e-6) to the float64
    direction (sqrt(1-z^2) cos 2pi v, sqrt(1-z^2) sin 2pi v, z) of the same draws."""
    n = 20000
    wf = _wave()
    geom = ds.FreeDiffusion()
    plan = oracle.oracle_plan(geom, D, wf.dt)
    plan["step_len"] = 1.0
    res = oracle.run_oracle(geom, D, wf, n, seed=77, plan=plan, r0=np.zeros((n, 3)))
    d = res["final_pos"].astype(np.float64)
    np.testing.assert_allclose(np.linalg.norm(d, axis=1), 1.0, atol=2e-7)
    out = np.array([oracle.philox4x32_10((w, 0, 0, 0), (77, 0)) for w in range(200)])
    u = (out[:, 0] >> 8) * 2.0 ** -24
    v = (out[:, 1] >> 8) * 2.0 ** -24
    z = 1 - 2 * u
    s = np.sqrt(1 - z * z)
    exact = np.stack([s * np.cos(2 * np.pi * v), s * np.sin(2 * np.pi * v), z], axis=1)
    np.testing.assert_allclose(d[:200], exact, atol=1e-6)
    # isotropy: mean 0, second moments 1/3 (3-sigma bands)
    assert np.all(np.abs(d.mean(axis=0)) < 3 * np.sqrt(1 / 3 / n))
    np.testing.assert_allclose((d ** 2).mean(axis=0), 1 / 3, atol=3 * np.sqrt(4 / 45 / n))


def test_free_step_is_fixed_length_unit_direction(oracle):
    """r_new = r + (n/|n|) * sqrt(6 D dt) (physics.py:299-301)."""
    res, plan = _one_step(oracle, ds.FreeDiffusion(), [1e-6, 2e-6, -3e-6], [0.3, -1.2, 2.0])
    n = np.array([0.3, -1.2, 2.0])
    expect = np.array([1e-6, 2e-6, -3e-6]) + n / np.linalg.norm(n) * np.sqrt(6 * D * 1e-4)
    np.testing.assert_allclose(res["final_pos"][0], expect, rtol=2e-6)


def test_cylinder_no_hit_is_straight(oracle):
    geom = ds.Cylinder(5e-6, (0, 0, 1))
    res, plan = _one_step(oracle, geom, [0.0, 0.0, 0.0], [1.0, 0.5, 0.2])
    n = np.array([1.0, 0.5, 0.2], np.float32)
    step = (n / np.float32(np.sqrt(np.float32((n[0] * n[0] + n[1] * n[1]) + n[2] * n[2])))) \
        * np.float32(plan["step_len"])
    np.testing.assert_array_equal(res["final_pos"][0], step.astype(np.float32))


@pytest.mark.parametrize("kind", ["cyl", "sph"])
def test_normal_incidence_mirror_and_bt_weight(oracle, kind):
    """Walker at distance delta inside the wall moving straight out by s > delta:
    mirror image 2R - x0 - s, and Brownstein-Tarr dlog = -2 (rho/D) (s - delta)
    with d_perp = (remaining - d) cos(alpha), cos(alpha) = 1 at normal incidence."""
    R, rho = 5e-6, 1e-5
    geom = (ds.Cylinder(R, (0, 0, 1), surface_relaxivity_t2=rho) if kind == "cyl"
            else ds.Sphere(R, surface_relaxivity_t2=rho))
    wf = _wave()
    plan = oracle.oracle_plan(geom, D, wf.dt)
    s = plan["step_len"]
    x0 = R - 0.4 * s
    res, plan = _one_step(oracle, geom, [x0, 0.0, 0.0], [1.0, 0.0, 0.0])
    x = float(res["final_pos"][0, 0])
    assert abs(x - (2 * R - x0 - s)) < 1e-4 * R * 1e-2 + 1e-12
    assert abs(float(res["final_pos"][0, 1])) == 0.0
    expect = -2.0 * (rho / D) * (s - (R - x0))
    assert abs(float(res["log_w"][0]) - expect) < 2e-5 * abs(expect) + 1e-9


def test_box1d_fold_and_overshoot(oracle):
    L, rho = 6e-6, 2e-5
    geom = ds.Box1D(L, surface_relaxivity_t2=rho)
    wf = _wave()
    s = oracle.oracle_plan(geom, D, wf.dt)["step_len"]
    for x0, sgn in [(L - 0.3 * s, 1.0), (0.2 * s, -1.0), (0.5 * L, 1.0)]:
        res, _ = _one_step(oracle, geom, [x0, 0.0, 0.0], [sgn, 0.0, 0.0])
        xr = x0 + sgn * s
        over = max(xr - L, 0.0) + max(-xr, 0.0)
        folded = xr % (2 * L)
        folded = 2 * L - folded if folded > L else folded
        assert abs(float(res["final_pos"][0, 0]) - folded) < 1e-12
        np.testing.assert_allclose(float(res["log_w"][0]), -2 * rho / D * over, rtol=1e-5, atol=1e-12)


@pytest.mark.parametrize("u,transmit", [(0.0, True), (0.9999999, False)])
def test_powles_decision(oracle, u, transmit):
    """Transmit iff u < min(1, 2 kappa d_perp / D); no weight on transmission."""
    R = 5e-6
    geom = ds.Cylinder(R, (0, 0, 1), permeability=1e-5, surface_relaxivity_t2=1e-5)
    wf = _wave()
    s = oracle.oracle_plan(geom, D, wf.dt)["step_len"]
    x0 = R - 0.5 * s
    res, plan = _one_step(oracle, geom, [x0, 0.0, 0.0], [1.0, 0.0, 0.0], u=u)
    if transmit:
        assert res["n_cross"][0] == 1 and res["comp_final"][0] == 1
        np.testing.assert_allclose(res["final_pos"][0, 0], x0 + s, rtol=1e-6)
        assert res["log_w"][0] == 0.0
    else:
        assert res["n_cross"][0] == 0 and res["comp_final"][0] == 0
        assert res["final_pos"][0, 0] < R
        assert res["log_w"][0] < 0.0


def _ray_circle_hits(p, d, centers, radii, L, tmax):
    """float64 brute force over periodic images: sorted (t, k, disc) entry roots
    along a ray segment of length tmax."""
    hits = []
    for k, (c, R) in enumerate(zip(centers, radii)):
        for ix in (-1, 0, 1):
            for iy in (-1, 0, 1):
                q = p - (c + L * np.array([ix, iy]))
                b = np.dot(d, q)
                disc = b * b - (np.dot(q, q) - R * R)
                if disc <= 0:
                    continue
                t = -b - np.sqrt(disc)
                if 0 < t < tmax:
                    hits.append((t, k, disc))
    return sorted(hits)


def test_packed_cylinders_hit_matches_float64_geometry(oracle):
    """Reflection off a packing: the float32 oracle matches a float64 ray/circle
    solution of the same step (hit point, specular direction, nudge) to 1e-10 m."""
    geom = packed_cyl()
    wf = _wave()
    plan = oracle.oracle_plan(geom, D, wf.dt)
    plan["step_len"] = 1.5e-6
    rng = np.random.default_rng(0)
    L = geom._L_float
    checked = 0
    for trial in range(400):
        r0 = geom.init_positions(1, trial)[0].astype(np.float64)
        nrm = rng.standard_normal(3)
        u = nrm / np.linalg.norm(nrm)
        dxy = u[:2] / np.linalg.norm(u[:2])
        sxy = np.linalg.norm(u[:2]) * plan["step_len"]
        hits = _ray_circle_hits(r0[:2], dxy, geom._centers_np, geom._radii_np, L, sxy)
        if not hits:
            continue
        t, k, disc = hits[0]
        # skip float32/float64 near-ties: grazing rays, ends of the segment, two walls
        if disc < 1e-16 or sxy - t < 1e-9 or (len(hits) > 1 and hits[1][0] - t < 1e-9):
            continue
        hit = r0[:2] + t * dxy
        c = geom._centers_np[k]
        q = hit - c
        q -= L * np.round(q / L)
        n = q / np.linalg.norm(q)
        refl = dxy - 2 * np.dot(dxy, n) * n
        nudge = plan["nudge"]
        expect = hit + nudge * n + refl * max(sxy - t - nudge, 0.0)
        qe = expect - geom._centers_np
        qe -= L * np.round(qe / L)
        if np.any(np.linalg.norm(qe, axis=1) < geom._radii_np * (1 + 1e-6)):
            continue   # the safety clamp would act: not a single-reflection case
        res, _ = _one_step(oracle, geom, r0, nrm, wall_plan={"step_len": plan["step_len"]})
        np.testing.assert_allclose(res["final_pos"][0, :2], expect, atol=1e-10)
        checked += 1
    assert checked > 20


@pytest.mark.parametrize("factory", [lambda: ds.Cylinder(3e-6, (0, 0, 1)),
                                     lambda: ds.Sphere(3e-6),
                                     lambda: ds.Cylinder(3e-6, (0, 0, 1), surface_relaxivity_t2=2e-5)])
def test_containment_after_walk(oracle, factory):
    """max |r| < R (1 + 3e-4) after a long walk (tests/test_cylinder.py:113-151)."""
    geom = factory()
    wf = _wave(n_t=200, dt=2e-4)
    res = oracle.run_oracle(geom, D, wf, 2000, seed=3)
    r = res["final_pos"].astype(np.float64)
    rad = np.linalg.norm(r[:, :2] if isinstance(geom, ds.Cylinder) else r, axis=1)
    assert rad.max() < geom.radius * (1 + 3e-4)
    assert rad.max() > 0.95 * geom.radius
    assert np.all(res["comp_final"] == 0)


@pytest.mark.parametrize("factory", [packed_cyl, packed_sph])
def test_packed_exterior_stays_outside(oracle, factory):
    geom = factory()
    wf = _wave(n_t=100, dt=2e-4)
    res = oracle.run_oracle(geom, D, wf, 1000, seed=4)
    assert np.all(res["comp_final"] == 0)
    L = geom._L_float
    dim = geom._centers_np.shape[1]
    p = res["final_pos"][:, :dim].astype(np.float64)
    for c, R in zip(geom._centers_np, geom._radii_np):
        q = p - c
        q -= L * np.round(q / L)
        assert np.min(np.linalg.norm(q, axis=1)) > R * (1 - 1e-4)


def test_free_pgse_matches_stejskal_tanner(oracle):
    """README quick-start workload: exp(-bD) within 3 Monte-Carlo SE."""
    wf = ds.set_b(ds.pgse(0.01, 0.04, 0.05, [[1, 0, 0]] * 3, 300), [0, 1e9, 2e9])
    N = 50000
    res = oracle.run_oracle(ds.FreeDiffusion(), D, wf, N, seed=0, threads=8)
    b = np.array([0, 1e9, 2e9])
    E = np.exp(-b * D)
    se = np.sqrt(np.maximum(((1 + np.exp(-4 * b * D)) / 2 - E ** 2) / N, 1e-12))
    assert np.all(np.abs(res["signal"] - E) <= 3 * se + 1e-9), (res["signal"], E, se)


def test_t2_decay_uses_n_t_samples(oracle):
    """Simulated time is n_t*dt (all samples including G[0], tests/test_t2.py:47-49)."""
    wf = _wave(n_t=500, dt=8e-5)
    T2 = 0.08
    res = oracle.run_oracle(ds.FreeDiffusion(), D, wf, 64, T2=T2)
    np.testing.assert_allclose(res["signal"][0], np.exp(-500 * 8e-5 / T2), rtol=2e-5)


def test_brownstein_tarr_fast_diffusion_sphere(oracle):
    """rho R / D << 1: S(t) = exp(-rho (S/V) t), S/V = 3/R, within 5 % in the rate
    (tests/physics/test_surface_relaxivity.py:127-230)."""
    R, rho = 2e-6, 1e-5
    geom = ds.Sphere(R, surface_relaxivity_t2=rho)
    wf = _wave(n_t=100, dt=2e-4)
    res = oracle.run_oracle(geom, D, wf, 4000, seed=8, threads=8)
    t = 100 * 2e-4
    rate = -np.log(res["signal"][0]) / t
    assert abs(rate / (3 * rho / R) - 1) < 0.05


def test_permeable_weights_only_on_reflection(oracle):
    """With rho set and kappa very large almost every wall event transmits, and a
    transmission carries no Brownstein-Tarr weight (geometries.py:419-422)."""
    geom = ds.Sphere(3e-6, permeability=1e3, surface_relaxivity_t2=1e-5)
    wf = _wave(n_t=20, dt=1e-4)
    res = oracle.run_oracle(geom, D, wf, 500, seed=2)
    assert res["n_cross"].sum() > 100
    assert np.mean(res["log_w"] != 0.0) < 0.01


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_ellipsoid_normal_incidence_along_each_axis(oracle, axis):
    """At a semi-axis end the outward normal is that axis: a walker moving straight
    out by s beyond the wall lands at 2c - x0 - s, with d_perp = s - (c - x0)."""
    semi = (3e-6, 2e-6, 4e-6)
    rho = 1e-5
    geom = ds.Ellipsoid(semi, surface_relaxivity_t2=rho)
    s = oracle.oracle_plan(geom, D, _wave().dt)["step_len"]
    c = semi[axis]
    x0 = c - 0.4 * s
    r0 = [0.0, 0.0, 0.0]
    r0[axis] = x0
    n = [0.0, 0.0, 0.0]
    n[axis] = 1.0
    res, _ = _one_step(oracle, geom, r0, n)
    x = float(res["final_pos"][0, axis])
    assert abs(x - (2 * c - x0 - s)) < 1e-6 * c + 1e-12      # float32 rounding
    assert np.all(res["final_pos"][0, [k for k in range(3) if k != axis]] == 0.0)
    expect = -2.0 * (rho / D) * (s - (c - x0))
    assert abs(float(res["log_w"][0]) - expect) < 2e-5 * abs(expect) + 1e-9


def test_ellipsoid_and_shell_containment(oracle):
    """Random walks stay inside a reflecting ellipsoid and inside the shell's
    outer wall (r < R_out) at every sample."""
    wf = _wave(n_t=40)
    for geom in (ds.Ellipsoid((3e-6, 1.5e-6, 2e-6)),
                 ds.PermeableShell(1.5e-6, 3e-6, 1e-4),
                 ds.PermeableShell(1.5e-6, 3e-6, 1e-4, kind="cylinder")):
        res = oracle.run_oracle(geom, D, wf, 400, seed=3, per_sample=True)
        pos = res["pos_full"].astype(np.float64)
        if isinstance(geom, ds.Ellipsoid):
            q = np.sum((pos / geom.semiaxes) ** 2, axis=-1)
            assert q.max() < 1.0
        else:
            rad = pos if geom.kind == "sphere" else pos[..., :2]
            assert np.linalg.norm(rad, axis=-1).max() < geom.r_outer
            assert res["n_hits"] > 0


@pytest.mark.parametrize("u,transmit", [(0.0, True), (0.9999999, False)])
def test_slab_membrane_decision_and_fold(oracle, u, transmit):
    """Crossing x = L/2: transmitted straight through, or mirrored to L - x_new with
    dlog = -2 (rho/D) |x_new - L/2|; walls at 0 and L fold (geometries.py:3270-3300)."""
    L, kappa, rho = 8e-6, 3e-5, 2e-5
    geom = ds.PermeableSlab1D(L, kappa, surface_relaxivity_t2=rho)
    s = oracle.oracle_plan(geom, D, _wave().dt)["step_len"]
    x0 = L / 2 - 0.3 * s
    res, _ = _one_step(oracle, geom, [x0, 1e-7, 0.0], [1.0, 0.0, 0.0], u=u)
    xn = np.float32(x0) + np.float32(s)
    if transmit:
        assert float(res["final_pos"][0, 0]) == float(np.float32(xn))
        assert res["n_cross"][0] == 1 and res["log_w"][0] == 0.0
    else:
        assert abs(float(res["final_pos"][0, 0]) - (L - float(xn))) < 1e-12
        expect = -2.0 * (rho / D) * (float(xn) - L / 2)
        assert abs(float(res["log_w"][0]) - expect) < 1e-5 * abs(expect)
    # outer wall: fold back inside [0, L]
    res2, _ = _one_step(oracle, geom, [0.2 * s, 0.0, 0.0], [-1.0, 0.0, 0.0], u=u)
    assert abs(float(res2["final_pos"][0, 0]) - 0.8 * s) < 1e-12


@pytest.mark.parametrize("kind", ["sphere", "cylinder"])
@pytest.mark.parametrize("u,transmit", [(0.0, True), (0.9999999, False)])
def test_shell_membrane_and_outer_wall(oracle, kind, u, transmit):
    """Outward through the membrane at R_in: transmitted straight, or mirrored to
    2 R_in - x0 - s (nudged back inside); at the outer wall always mirrored to
    2 R_out - x0 - s (geometries.py:3349-3426)."""
    Rin, Rout, kappa = 2e-6, 4e-6, 5e-5
    geom = ds.PermeableShell(Rin, Rout, kappa, kind=kind)
    s = oracle.oracle_plan(geom, D, _wave().dt)["step_len"]
    x0 = Rin - 0.3 * s
    res, _ = _one_step(oracle, geom, [x0, 0.0, 0.0], [1.0, 0.0, 0.0], u=u)
    x = float(res["final_pos"][0, 0])
    if transmit:
        assert abs(x - (x0 + s)) < 1e-12 and res["n_cross"][0] == 1
    else:
        assert abs(x - (2 * Rin - x0 - s)) < 2e-4 * Rin and x < Rin
    x1 = Rout - 0.3 * s
    res2, _ = _one_step(oracle, geom, [x1, 0.0, 0.0], [1.0, 0.0, 0.0], u=u)
    assert abs(float(res2["final_pos"][0, 0]) - (2 * Rout - x1 - s)) < 2e-4 * Rout
    assert res2["n_cross"][0] == 0


@pytest.mark.parametrize("u,transmit", [(0.0, True), (0.9999999, False)])
def test_producer_occupancy_rule(oracle, u, transmit):
    """Occupancy of the producer's compartment-1 channel (dmipy_sim/core.py:1166-1185):
    through a membrane it follows the transmit decision; it keys on the geometry's
    radius, so a slab walker with x < L/2 but |r| >= L/2, and a cylindrical-shell
    walker inside R_in radially but with |r| >= R_in, count as outside; an Ellipsoid
    (no radius) always reads 0."""
    L, kappa = 8e-6, 3e-5
    slab = ds.PermeableSlab1D(L, kappa)
    s = oracle.oracle_plan(slab, D, _wave().dt)["step_len"]
    res, _ = _one_step(oracle, slab, [L / 2 - 0.3 * s, 0.0, 0.0], [1.0, 0.0, 0.0], u=u,
                       per_sample=True)
    assert res["occ_full"][0, 0] == (1.0 if transmit else 0.0)
    assert res["occ_full"][0, 0] == res["comp_full"][0, 0]
    res, _ = _one_step(oracle, slab, [0.1 * L, 0.6 * L, 0.0], [0.0, 0.0, 1.0], u=u,
                       per_sample=True)
    assert res["comp_full"][0, 0] == 0 and res["occ_full"][0, 0] == 1.0
    Rin = 2e-6
    for kind in ("sphere", "cylinder"):
        shell = ds.PermeableShell(Rin, 2 * Rin, kappa, kind=kind)
        s = oracle.oracle_plan(shell, D, _wave().dt)["step_len"]
        res, _ = _one_step(oracle, shell, [Rin - 0.3 * s, 0.0, 0.0], [1.0, 0.0, 0.0], u=u,
                           per_sample=True)
        assert res["occ_full"][0, 0] == (1.0 if transmit else 0.0)
    cyl_shell = ds.PermeableShell(Rin, 2 * Rin, kappa, kind="cylinder")
    res, _ = _one_step(oracle, cyl_shell, [0.5 * Rin, 0.0, 1.5 * Rin], [0.0, 1.0, 0.0], u=u,
                       per_sample=True)
    assert res["comp_full"][0, 0] == 0 and res["occ_full"][0, 0] == 1.0
    ell = ds.Ellipsoid((3e-6, 2e-6, 4e-6), permeability=kappa)
    res, _ = _one_step(oracle, ell, [0.0, 0.0, 3.9e-6], [0.0, 0.0, 1.0], u=0.0, per_sample=True)
    assert res["occ_full"][0, 0] == 0.0


def test_producer_occupancy_is_mean_of_sub_step_compartments(oracle):
    """With several sub-steps per sample the channel is the fraction of sub-steps in
    compartment 1; for a permeable Sphere the rule equals the walker's compartment,
    so a one-sub-step walk's per-sample compartments reproduce it exactly."""
    geom = ds.Sphere(3e-6, permeability=5e-4)
    wf = _wave(n_t=60, dt=2e-5)
    p = oracle.oracle_plan(geom, D, wf.dt)
    p.update(n_sub=1, dt_sub=float(np.float32(wf.dt)))
    res = oracle.run_oracle(geom, D, wf, 1000, seed=2, plan=p, per_sample=True)
    np.testing.assert_array_equal(res["occ_full"], res["comp_full"].astype(np.float32))
    assert 0 < res["occ_full"].mean() < 1
    p4 = dict(p, n_sub=4, step_len=float(np.float32(p["step_len"]) / np.float32(2.0)))
    res4 = oracle.run_oracle(geom, D, wf, 1000, seed=2, plan=p4, per_sample=True)
    occ = res4["occ_full"] * 4
    np.testing.assert_array_equal(occ, np.round(occ))


# ------------------------------------------------------------ MyelinatedCylinder
def _myel(**kw):
    base = dict(inner_radius=2e-6, outer_radius=3.5e-6, orientation=(0, 0, 1), D_intra=2e-9,
                D_extra=2e-9, D_myelin=1e-9)
    base.update(kw)
    return ds.MyelinatedCylinder(**base)


@pytest.mark.parametrize("u,transmit", [(0.0, True), (0.9999999, False)])
def test_myelin_membranes_transmit_or_mirror(oracle, u, transmit):
    """Crossing R_in outward from the axon and R_out inward from outside: transmitted
    straight on (compartment changes), or mirrored radially to 2R - |r_new|
    (physics.py:560-611)."""
    g = _myel(kappa_inner=1e-4, kappa_outer=1e-4)
    p = oracle.oracle_plan(g, None, _wave().dt)
    Ri, Ro = np.float32(g.inner_radius), np.float32(g.outer_radius)
    s0, s2 = np.float32(p["m_step"][0]), np.float32(p["m_step"][2])
    x0 = float(Ri - np.float32(0.3) * s0)
    res, _ = _one_step(oracle, g, [x0, 0.0, 0.0], [1.0, 0.0, 0.0], u=u)
    xn = np.float32(x0) + s0
    if transmit:
        assert res["final_pos"][0, 0] == xn and res["comp_final"][0] == 1 and res["n_cross"][0] == 1
    else:
        assert res["final_pos"][0, 0] == np.float32(2.0) * Ri - xn and res["comp_final"][0] == 0
    x1 = float(Ro + np.float32(0.3) * s2)
    res, _ = _one_step(oracle, g, [x1, 0.0, 0.0], [-1.0, 0.0, 0.0], u=u)
    xn = np.float32(x1) - s2
    if transmit:
        assert res["final_pos"][0, 0] == xn and res["comp_final"][0] == 1
    else:
        assert res["final_pos"][0, 0] == np.float32(2.0) * Ro - xn and res["comp_final"][0] == 2


def test_myelin_overshoot_mirror_then_clamp(oracle):
    """A step of 3.5 R_in from the axis mirrors to radius -1.5 R_in (through the axis),
    which the compartment-0 clamp pulls back to R_in - NUDGE (physics.py:621-631)."""
    Ri = 1e-6
    dt = _wave().dt
    D = (3.5 * Ri) ** 2 / (6.0 * dt)
    g = _myel(inner_radius=Ri, outer_radius=2e-6, D_intra=D)
    p = oracle.oracle_plan(g, None, dt)
    s0 = np.float32(p["m_step"][0])
    assert abs(float(s0) / Ri - 3.5) < 1e-6
    res, _ = _one_step(oracle, g, [0.0, 0.0, 0.0], [1.0, 0.0, 0.0], u=0.5)
    target = np.float32(Ri) - np.float32(p["nudge"])
    assert res["final_pos"][0, 0] == -target and res["final_pos"][0, 1] == 0.0
    assert res["comp_final"][0] == 0


def test_myelin_start_allotment_stuck_pool_and_weighted_t2(oracle):
    """In-kernel starts: compartment counts are round(N A_c / A) by cross-section
    area (extra-axonal bounded at 2 R_out); a D_myelin = 0 pool never moves; with
    impermeable membranes the b = 0 signal is the spin-weighted mixture
    sum_c n_c w_c exp(-n_t dt/T2_c) / sum_c n_c w_c (_ens, core.py:426-429)."""
    g = _myel(D_myelin=0.0, T2_intra=0.08, T2_myelin=0.012, T2_extra=0.1,
              water_fractions=(1.0, 0.4, 0.7))
    n_t, dt, N = 60, 2e-4, 3000
    wf = ds.Waveform(G=np.zeros((1, n_t, 3), np.float32), dt=dt, echo_idx=n_t - 1)
    res = oracle.run_oracle(g, None, wf, N, seed=4, per_sample=True)
    Ri, Ro = g.inner_radius, g.outer_radius
    A = np.array([Ri ** 2, Ro ** 2 - Ri ** 2, 4 * Ro ** 2 - Ro ** 2])
    n0, n1 = int(round(N * A[0] / A.sum())), int(round(N * A[1] / A.sum()))
    counts = np.array([n0, n1, N - n0 - n1])
    np.testing.assert_array_equal(np.bincount(res["comp_origin"], minlength=3), counts)
    assert np.all(res["comp_origin"][:n0] == 0) and np.all(res["comp_origin"][n0:n0 + n1] == 1)
    sheath = res["comp_origin"] == 1
    pos = res["pos_full"][sheath]
    assert np.all(pos == pos[:, :1, :])
    assert np.all(res["comp_final"] == res["comp_origin"])
    w = np.array([1.0, 0.4, 0.7])
    T2 = np.array([0.08, 0.012, 0.1])
    expect = np.sum(counts * w * np.exp(-n_t * dt / T2)) / np.sum(counts * w)
    assert abs(res["signal"][0] - expect) < 2e-6 * expect


# ------------------------------------------------------ PackedMyelinatedCylinders
def _pmyel(n=8, seed=1, **kw):
    rng = np.random.default_rng(0)
    ri, gr, c = ds.pack_myelinated_cylinders(rng.uniform(1e-6, 2e-6, n), 0.7, 0.5, seed=seed)
    L = float(np.sqrt(np.pi * np.sum((ri / gr) ** 2) / 0.5))
    return ds.PackedMyelinatedCylinders(ri, gr, c, L, N_max=16, **kw), L


def _axon_dist(g, L, pos):
    n = g.N_actual
    d = pos[..., None, :2].astype(np.float64) - g._centers_np[:n]
    d -= L * np.floor(d / L + 0.5)
    return np.sqrt(np.sum(d ** 2, -1))


def test_packed_myelin_allotment_pools_and_weighted_t2(oracle):
    """Start classes follow the area allotment in index order (intra of each axon,
    myelin of each axon, the rest extra-axonal); D_myelin = 0 water stays put up
    to the cell re-wrap's rounding; intra walkers never leave their axon; with
    impermeable walls the b = 0 signal is the spin-weighted mixture
    sum_c n_c w_c exp(-n_t dt/T2_c) / sum_c n_c w_c with w = (1, 0.4, 1)."""
    g, L = _pmyel(D_myelin=0.0, T2_intra=0.08, T2_myelin=0.015, T2_extra=0.1)
    n_t, dt, N = 100, 1e-4, 2000
    wf = ds.Waveform(G=np.zeros((1, n_t, 3), np.float32), dt=dt, echo_idx=n_t - 1)
    res = oracle.run_oracle(g, None, wf, N, seed=3, per_sample=True)
    n_in, n_my, n_ex = g.start_counts(N)
    co = res["comp_origin"]
    np.testing.assert_array_equal(np.bincount(co, minlength=3), [n_in.sum(), n_my.sum(), n_ex])
    assert np.all(co[:n_in.sum()] == 0) and np.all(co[n_in.sum():n_in.sum() + n_my.sum()] == 1)
    pos = res["pos_full"]
    assert np.abs(pos[co == 1] - pos[co == 1][:, :1]).max() < 1e-11
    dist = _axon_dist(g, L, pos[co == 0])
    assert np.all(np.min(dist - g._inner_radii_np[:g.N_actual], axis=-1) < 0)
    assert np.all(res["comp_final"] == co)
    cnt = np.bincount(co, minlength=3)
    w = np.array([1.0, 0.4, 1.0])
    T2 = np.array([0.08, 0.015, 0.1])
    expect = np.sum(cnt * w * np.exp(-n_t * dt / T2)) / np.sum(cnt * w)
    assert abs(res["signal"][0] - expect) < 1e-6 * expect


def test_packed_myelin_phase_uses_unwrapped_position(oracle):
    """With G nonzero at one sample only, phi = GAMMA dt G . r_unwrapped(t); the
    unwrapped position moves by at most three steps between samples (a step plus
    the exterior mirror, which moves a walker by twice its penetration), differs
    from the in-cell position by whole cells, and some walkers do cross the cell
    edge (physics.py:1040-1047)."""
    g, L = _pmyel(D_extra=3e-9)
    n_t, dt, N = 400, 1e-4, 1500
    G = np.zeros((4, n_t, 3), np.float32)
    G[0, -1, 0] = G[1, -2, 0] = 0.05
    G[2, -1, 1] = G[3, -2, 1] = 0.05
    wf = ds.Waveform(G=G, dt=dt, echo_idx=n_t - 1)
    p = oracle.oracle_plan(g, None, dt)
    res = oracle.run_oracle(g, None, wf, N, seed=5, per_sample=True, phase64=True)
    uw = res["phi"].astype(np.float64) / (p["gamma_dt_sub"] * 0.05)     # (N, 4)
    ux_t, ux_p, uy_t, uy_p = uw.T
    step = float(p["pm_step"][:, 2].max())
    assert np.all(np.hypot(ux_t - ux_p, uy_t - uy_p) <= 3.0 * step * 1.001 + 1e-12)
    pos = res["pos_full"][:, -1, :2].astype(np.float64)
    m = np.stack([ux_t - pos[:, 0], uy_t - pos[:, 1]], 1) / L
    assert np.all(np.abs(m - np.round(m)) < 1e-3)
    assert np.any(np.round(m) != 0)


def test_exact_phase_is_the_float64_sum_over_the_trajectory(oracle):
    """The oracle's exact phase (the FP64 reference the kernel's phase error is
    stated against) equals sum_t gamma dt_sub G(t) . r(t) summed in float64 over
    the walk's float32 positions: with FreeDiffusion n_sub = 1, so every sub-step
    position is a per-sample output.  Its signal is mean exp(log_w) cos(phi_x), and
    the float32 phase of the same walk stays within float32 accumulation error."""
    rng = np.random.default_rng(3)
    n_t, N, M = 60, 200, 5
    G = rng.uniform(-0.5, 0.5, (M, n_t, 3)).astype(np.float32)
    wf = ds.Waveform(G=G, dt=1e-4, echo_idx=n_t - 1)
    geom = ds.FreeDiffusion()
    p = oracle.oracle_plan(geom, D, wf.dt, T2=0.05)
    assert p["n_sub"] == 1
    res = oracle.run_oracle(geom, D, wf, N, seed=2, T2=0.05, per_sample=True, exact_phase=True)
    pos = res["pos_full"].astype(np.float64)                       # (N, n_t, 3)
    phi = p["gamma_dt_sub"] * np.einsum("mtc,wtc->wm", G.astype(np.float64), pos)
    assert np.abs(phi).max() > 0.5                                 # a non-trivial phase
    np.testing.assert_allclose(res["phi_x"], phi, rtol=1e-12, atol=1e-12)
    w = np.exp(res["log_w"].astype(np.float64))
    np.testing.assert_allclose(res["signal_x"], (w[:, None] * np.cos(phi)).mean(0), atol=1e-12)
    np.testing.assert_allclose(res["signal_x_im"], (w[:, None] * np.sin(phi)).mean(0), atol=1e-12)
    assert np.max(np.abs(res["phi"] - phi)) < 1e-5 * max(1.0, np.abs(phi).max())
    assert np.max(np.abs(res["phi"] - phi)) > 0.0                  # float32 does round


def test_first_walker_keys_the_streams(oracle):
    """A walk of walkers [k, k + n) with first_walker = k is the slice of the
    whole walk (the same Philox streams and start positions)."""
    geom = packed_cyl(kappa=5e-5)
    wf = ds.set_b(ds.pgse(0.002, 0.006, 0.05, [[1, 0, 0], [0, 1, 0]], 12, slew_rate=np.inf),
                  [1e9, 2e9])
    whole = oracle.run_oracle(geom, D, wf, 96, seed=4)
    part = oracle.run_oracle(geom, D, wf, 40, seed=4, first_walker=50)
    np.testing.assert_array_equal(part["final_pos"], whole["final_pos"][50:90])
    np.testing.assert_array_equal(part["log_w"], whole["log_w"][50:90])
    np.testing.assert_array_equal(part["n_cross"], whole["n_cross"][50:90])
    np.testing.assert_array_equal(part["phi"], whole["phi"][50:90])
