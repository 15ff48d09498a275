"""GPU walk against closed-form physics (the reference's physics specification,
tests/ of dmipy-sim) and against the oracle with independent random streams."""
import numpy as np
import pytest
from scipy.special import j1
from scipy.optimize import curve_fit

import dmipy_sim_amd as ds
from dmipy_sim_amd import configs
from substrates import D, small_pgse, packed_cyl, packed_sph

pytestmark = pytest.mark.gpu


def mc_se(E_cos2, E, n):
    return np.sqrt(np.maximum((1 + E_cos2) / 2 - E ** 2, 1e-12) / n)


def test_free_pgse_readme_workload(gpu):
    """50k walkers, b = 0, 1e9, 2e9: exp(-bD) within 3 Monte-Carlo SE (README.md:36-41)."""
    cfg = configs.free_pgse()
    b = np.array([0, 1e9, 2e9])
    S = gpu.simulate(cfg["n_walkers"], D, cfg["waveform"], cfg["geometry"], seed=0)
    assert S.dtype == np.float32 and S.shape == (3,)
    E = np.exp(-b * D)
    assert np.all(np.abs(S - E) <= 3 * mc_se(np.exp(-4 * b * D), E, cfg["n_walkers"]) + 1e-7), S


def test_free_1d_hundred_b_values(gpu):
    """tests/test_free_1d.py:14-35: 100 b-values to 2e9, atol 0.01."""
    b = np.linspace(1, 2e9, 100)
    wf = ds.set_b(ds.pgse(8e-3, 50e-3, 1.0, np.tile([1.0, 0, 0], (100, 1)), 1000), b)
    S = gpu.simulate(100_000, D, wf, ds.FreeDiffusion(), seed=123)
    np.testing.assert_allclose(S, np.exp(-b * D), atol=0.01)


def test_free_3d_directions_and_ogse(gpu):
    """tests/test_free_3d.py, tests/test_free_ogse.py: isotropy and OGSE exp(-bD)."""
    dirs = np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 0], [1, 0, 1], [0, 1, 1]], float)
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    wf = ds.set_b(ds.pgse(10e-3, 30e-3, 1.0, dirs, 500), np.full(6, 1e9))
    S = gpu.simulate(100_000, D, wf, ds.FreeDiffusion(), seed=5)
    np.testing.assert_allclose(S, np.exp(-1e9 * D), atol=0.01)
    og = ds.set_b(ds.ogse(50.0, 0.04, 0.05, [[1, 0, 0], [0, 0, 1]], 801), [5e8, 1e9])
    S = gpu.simulate(100_000, D, og, ds.FreeDiffusion(), seed=6)
    np.testing.assert_allclose(S, np.exp(-np.array([5e8, 1e9]) * D), atol=0.01)


def test_t2_exact_and_scaling(gpu):
    """b = 0: exp(-n_t dt / T2) to float32 precision (tests/test_t2.py:24-35);
    T2 multiplies the diffusion signal (tests/test_t2.py:38-63)."""
    wf0 = ds.set_b(ds.pgse(1e-3, 40e-3, 1.0, [[1, 0, 0]], 500), [1.0])
    S = gpu.simulate(20_000, D, wf0, ds.FreeDiffusion(), seed=1, T2=0.08)
    assert abs(S[0] - np.exp(-500 * wf0.dt / 0.08)) < 2e-5
    b = np.linspace(1e8, 2e9, 20)
    wf = ds.set_b(ds.pgse(1e-3, 40e-3, 1.0, np.tile([1, 0, 0.], (20, 1)), 500), b)
    n = 100_000
    s0 = gpu.simulate(n, D, wf, ds.FreeDiffusion(), seed=2)
    s1 = gpu.simulate(n, D, wf, ds.FreeDiffusion(), seed=2, T2=0.06)
    np.testing.assert_allclose(s1, s0 * np.exp(-500 * wf.dt / 0.06), atol=3 / np.sqrt(n))
    s2 = gpu.simulate(n, D, wf, ds.FreeDiffusion(), seed=2, T2=None)
    np.testing.assert_array_equal(s0, s2)


def test_cpmg_pure_t2_and_monotone(gpu):
    """tests/test_cpmg_multiecho.py:12-29."""
    TE, T2, n = 12e-3, 60e-3, 8
    wf = ds.cpmg(n, TE, 0.0, [[1, 0, 0]], n_t_per_echo=60)
    S = np.asarray(gpu.simulate_cpmg(20_000, D, wf, ds.Sphere(5e-6), T2=T2, seed=1)).ravel()
    assert S.shape == (n,)
    np.testing.assert_allclose(S, np.exp(-np.arange(1, n + 1) * TE / T2), atol=0.01)
    wf = ds.cpmg(6, 10e-3, 0.02, [[1, 0, 0]], n_t_per_echo=50)
    S = np.asarray(gpu.simulate_cpmg(20_000, D, wf, ds.FreeDiffusion(), T2=80e-3, seed=2)).ravel()
    assert S.shape == (6,) and np.all(np.diff(S) <= 1e-3)


def test_cylinder_parallel_free_perpendicular_restricted(gpu):
    """tests/test_cylinder.py:80-201: parallel gradient gives free diffusion,
    perpendicular signal exceeds free, orientation sign does not matter."""
    b = np.array([5e8, 1e9, 2e9])
    geom = ds.Cylinder(5e-6, (0, 0, 1))
    wf_par = ds.set_b(ds.pgse(10e-3, 30e-3, 1.0, np.tile([0, 0, 1.], (3, 1)), 400), b)
    wf_perp = ds.set_b(ds.pgse(10e-3, 30e-3, 1.0, np.tile([1, 0, 0.], (3, 1)), 400), b)
    n = 200_000
    Sp = gpu.simulate(n, D, wf_par, geom, seed=3)
    np.testing.assert_allclose(Sp, np.exp(-b * D), atol=0.01)
    Sx = gpu.simulate(n, D, wf_perp, geom, seed=3)
    assert np.all(Sx > np.exp(-b * D) + 0.05)
    Sneg = gpu.simulate(n, D, wf_perp, ds.Cylinder(5e-6, (0, 0, -1)), seed=3)
    np.testing.assert_allclose(Sneg, Sx, atol=3 / np.sqrt(n))


def _sgp(pore, q_first_zero, E_fn, geom, bvec, n=1_000_000):
    """Short-pulse, long-time q-sweep (tests/physics/test_sgp_pore_sizes.py:59-123):
    delta = 1e-3 tau, DELTA = 5 tau, tau = pore^2 / D; 64 q in [0, 3 q_zero];
    chi^2/dof < 2 with sigma = 1/sqrt(N), fitted size within 2 %."""
    tau = pore ** 2 / D
    delta, DELTA = 1e-3 * tau, 5 * tau
    n_t = int((DELTA + delta) / (delta / 2)) + 2
    q = np.linspace(0, 3 * q_first_zero, 64)
    tmpl = ds.pgse(delta, DELTA, 1.0, np.tile(bvec, (64, 1)), n_t)
    bt = ds.calc_b(tmpl)
    b_target = q ** 2 * (DELTA - delta / 3)
    scale = np.sqrt(np.where(b_target > 0, b_target / bt, 0.0)).astype(np.float32)
    wf = ds.Waveform(G=(tmpl.G * scale[:, None, None]).astype(np.float32), dt=tmpl.dt,
                     echo_idx=tmpl.echo_idx)
    E = np.asarray(ds.simulate(n, D, wf, geom, seed=42), np.float64)
    chi2 = np.sum((E - E_fn(q, pore)) ** 2 * n) / len(q)
    popt, _ = curve_fit(E_fn, q[1:], E[1:], p0=[pore], bounds=(pore * 0.5, pore * 2))
    return chi2, popt[0]


def _e_slab(q, d):
    x = q * d
    return np.where(np.abs(x) < 1e-12, 1.0, (2 * np.sin(x / 2) / np.where(x == 0, 1, x)) ** 2)


def _e_cyl(q, R):
    x = q * R
    return np.where(x < 1e-12, 1.0, (2 * j1(x) / np.where(x == 0, 1, x)) ** 2)


def _e_sph(q, R):
    x = np.where(q * R < 1e-12, 1.0, q * R)
    j = (np.sin(x) - x * np.cos(x)) / x ** 2
    return np.where(q * R < 1e-12, 1.0, (3 * j / x) ** 2)


@pytest.mark.parametrize("kind", ["slab", "cylinder", "sphere"])
def test_sgp_form_factors(gpu, kind):
    if kind == "slab":
        d = 10e-6
        chi2, fit = _sgp(d, 2 * np.pi / d, _e_slab, ds.Box1D(d), [1.0, 0, 0])
    elif kind == "cylinder":
        d = 5e-6
        chi2, fit = _sgp(d, 3.8317 / d, _e_cyl, ds.Cylinder(d, (0, 0, 1)), [1.0, 0, 0])
    else:
        d = 6e-6
        chi2, fit = _sgp(d, 4.4934 / d, _e_sph, ds.Sphere(d), [1.0, 0, 0])
    assert chi2 < 2.0, chi2
    assert abs(fit / d - 1) < 0.02, fit


@pytest.mark.parametrize("kind,sv", [("box", lambda L: 2 / L), ("cyl", lambda R: 2 / R),
                                     ("sph", lambda R: 3 / R)])
def test_brownstein_tarr_fast_diffusion(gpu, kind, sv):
    """rho a / D << 1: S(t) = exp(-rho (S/V) t) within 5 % in the rate
    (tests/physics/test_surface_relaxivity.py:127-361)."""
    a, rho = 4e-6, 1e-5
    geom = {"box": ds.Box1D(a, surface_relaxivity_t2=rho),
            "cyl": ds.Cylinder(a, (0, 0, 1), surface_relaxivity_t2=rho),
            "sph": ds.Sphere(a, surface_relaxivity_t2=rho)}[kind]
    n_t, T = 400, 0.08
    wf = ds.Waveform(G=np.zeros((1, n_t, 3), np.float32), dt=T / n_t, echo_idx=n_t - 1)
    S = gpu.simulate(200_000, D, wf, geom, seed=9)[0]
    rate = -np.log(S) / T
    assert abs(rate / (rho * sv(a)) - 1) < 0.05, rate / (rho * sv(a))


def test_surface_relaxation_scales_with_surface_to_volume(gpu):
    """Halving the radius doubles the relaxation rate (S/V slope 1)."""
    rho = 1e-5
    n_t, T = 400, 0.06
    wf = ds.Waveform(G=np.zeros((1, n_t, 3), np.float32), dt=T / n_t, echo_idx=n_t - 1)
    r = []
    for R in (3e-6, 6e-6):
        S = gpu.simulate(200_000, D, wf, ds.Sphere(R, surface_relaxivity_t2=rho), seed=4)[0]
        r.append(-np.log(S) / T)
    assert abs(r[0] / r[1] - 2.0) < 0.1


def test_permeability_compartment_exchange(gpu):
    """tests/test_compartment_tagging.py:190-257: with kappa > 0 more than 5 % of
    walkers change compartment; impermeable walls keep every walker (0 %)."""
    wf = small_pgse(n_t=100, T=0.05)
    perm = packed_cyl(kappa=1e-4)
    _, pos, c0, cf = gpu.simulate(20_000, D, wf, perm, seed=3, return_positions=True,
                                  return_compartments='final')
    assert np.mean(c0 != cf) > 0.05
    imp = packed_cyl()
    _, c0, cf = gpu.simulate(20_000, D, wf, imp, seed=3, return_compartments='final')
    assert np.all(c0 == 0) and np.all(cf == 0)


def test_high_permeability_is_free_diffusion(gpu):
    """tests/physics/test_permeability.py:364-420: kappa -> large makes walls
    transparent and the signal free (exp(-bD))."""
    b = np.array([5e8, 1e9, 2e9])
    wf = ds.set_b(ds.pgse(5e-3, 20e-3, 1.0, np.tile([1, 0, 0.], (3, 1)), 200), b)
    S = gpu.simulate(100_000, D, wf, ds.Sphere(3e-6, permeability=10.0), seed=5)
    np.testing.assert_allclose(S, np.exp(-b * D), atol=0.015)


def test_permeability_monotone(gpu):
    """Signal perpendicular to packed cylinders falls as kappa grows
    (tests/test_packed_cylinders_permeability.py:90-206)."""
    wf = ds.set_b(ds.pgse(5e-3, 25e-3, 1.0, [[1, 0, 0.]], 150), [2e9])
    S = [gpu.simulate(100_000, D, wf, packed_cyl(kappa=k), seed=7)[0] for k in (1e-6, 1e-4, 1e-3)]
    assert S[0] > S[1] > S[2]


def test_stimulated_echo_half_amplitude(gpu):
    wf = small_pgse(n_t=40)
    wf2 = ds.Waveform(G=wf.G, dt=wf.dt, echo_idx=wf.echo_idx, stimulated_echo=True)
    a = gpu.simulate(5000, D, wf, ds.Sphere(3e-6), seed=1)
    b = gpu.simulate(5000, D, wf2, ds.Sphere(3e-6), seed=1)
    np.testing.assert_array_equal(b, a * np.float32(0.5))


def test_pgste_t1_and_stimulated_echo(gpu):
    """tests/test_pgste_t1.py: a PGSTE on free diffusion gives 0.5 exp(-bD)
    exp(-TM/T1) (T1 acts only while stored), 0.5 exp(-bD) without T1, the signal
    falls with the mixing time, and a PGSE (always transverse) is unchanged by T1
    bit for bit.  3 Monte-Carlo SE."""
    N, b = 40_000, 1e9

    def pgste_at(TM):
        return ds.set_b(ds.pgste(5e-3, TM, 1.0, [[1.0, 0.0, 0.0]], 600), [b])

    wf = pgste_at(40e-3)
    b_eff = ds.calc_b(wf)[0]
    E = np.exp(-b_eff * D)
    se = 3 * mc_se(np.exp(-4 * b_eff * D), E, N) * 0.5 + 1e-6
    S1 = gpu.simulate(N, D, wf, ds.FreeDiffusion(), seed=3, T1=1.0)[0]
    assert abs(S1 - 0.5 * E * np.exp(-40e-3 / 1.0)) <= se, S1
    S0 = gpu.simulate(N, D, wf, ds.FreeDiffusion(), seed=3)[0]
    assert abs(S0 - 0.5 * E) <= se, S0
    Ss = [gpu.simulate(N, D, pgste_at(TM), ds.FreeDiffusion(), seed=3, T1=0.8)[0]
          for TM in (10e-3, 40e-3, 80e-3)]
    assert Ss[0] >= Ss[1] - 1e-3 and Ss[1] >= Ss[2] - 1e-3
    pg = ds.set_b(ds.pgse(5e-3, 45e-3, 1.0, [[1.0, 0.0, 0.0]], 600), [b])
    np.testing.assert_array_equal(gpu.simulate(N, D, pg, ds.FreeDiffusion(), seed=3),
                                  gpu.simulate(N, D, pg, ds.FreeDiffusion(), seed=3, T1=1.0))


def test_tensor_encodings_on_free_diffusion(gpu):
    """tests/test_btensor.py MC checks: for isotropic free diffusion the signal is
    exp(-trace(B) D) whatever the B-tensor's shape: linear, spherical and planar
    encodings at the same b agree with exp(-bD) within 3 SE."""
    N, b = 60_000, 1.5e9
    wfs = [ds.set_b(ds.pgse(10e-3, 40e-3, 0.06, [[0.0, 0.0, 1.0]], 1000), [b]),
           ds.set_b(ds.ste(10e-3, 40e-3, 0.06, 1000), [b]),
           ds.set_b(ds.pte(10e-3, 40e-3, 0.06, [0.0, 0.0, 1.0], 1000), [b])]
    E = np.exp(-b * D)
    for wf in wfs:
        S = gpu.simulate(N, D, wf, ds.FreeDiffusion(), seed=5)[0]
        assert abs(S - E) <= 3 * mc_se(np.exp(-4 * b * D), E, N) + 1e-6, (S, E)


def test_outputs_shapes_and_batches(gpu):
    """simulate() return tuple shapes (core.py:737-762) and walker batching:
    chunked run agrees with the single run statistically (core.py:765-818)."""
    wf = small_pgse(n_t=50)
    geom = packed_sph(kappa=5e-5)
    n = 3000
    out = gpu.simulate(n, D, wf, geom, seed=2, return_positions='full',
                       return_compartments='full', return_walker_signals=True)
    sig, pos, c0, cfull, lw, phi = out
    assert sig.shape == (16,) and pos.shape == (n, 50, 3) and cfull.shape == (n, 50)
    assert c0.shape == (n,) and lw.shape == (n,) and phi.shape == (n, 16)
    recon = np.mean(np.exp(lw)[:, None] * np.cos(phi.astype(np.float64)), axis=0)
    np.testing.assert_allclose(recon, sig, atol=2e-5)
    sig2, cf = gpu.simulate(n, D, wf, geom, seed=2, return_compartments='final')[::2]
    np.testing.assert_array_equal(cf, cfull[:, -1])
    big = gpu.simulate(40_000, D, wf, ds.Sphere(3e-6), seed=4)
    chunked = gpu.simulate(40_000, D, wf, ds.Sphere(3e-6), seed=4, walker_batch_size=15_000)
    np.testing.assert_allclose(chunked, big, atol=4 / np.sqrt(40_000))
    mix = gpu.simulate_mixture([dict(fraction=0.3, n_walkers=10_000, diffusivity=D,
                                     geometry=ds.Sphere(3e-6)),
                                dict(fraction=0.7, n_walkers=10_000, diffusivity=D,
                                     geometry=ds.FreeDiffusion())], wf, seed=3)
    assert mix.shape == (16,)


def test_independent_streams_agree_with_oracle(gpu, oracle):
    """GPU (Philox seed 1) vs oracle (injected numpy normals) on the same
    relaxing, permeable packing: within 3 combined Monte-Carlo SE."""
    wf = small_pgse(n_t=40)
    geom = packed_cyl(kappa=5e-5, rho=1e-5)
    n = 20_000
    g = gpu.walk(geom, D, wf, n, seed=1)["signal"]
    plan = oracle.oracle_plan(geom, D, wf.dt)
    rng = np.random.default_rng(12)
    S = wf.G.shape[1] * plan["n_sub"]
    r0 = geom.init_positions(n, 5)
    o = oracle.run_oracle(geom, D, wf, n, r0=r0, threads=8,
                          inject_normals=rng.standard_normal((n, S, 3)).astype(np.float32),
                          inject_uniforms=rng.random((n, S)).astype(np.float32))["signal"]
    se = np.sqrt(2.0 / n)
    assert np.all(np.abs(g - o) <= 3 * se), np.max(np.abs(g - o)) / se


def test_autotune_picks_a_timed_block_and_walk_uses_it(gpu):
    from dmipy_sim_amd import autotune
    autotune.clear()
    wf = small_pgse(n_t=40, n_dirs=8)               # 32 measurements
    geom = gpu.Cylinder(3e-6, (0, 0, 1), surface_relaxivity_t2=1e-5)
    best, ms = autotune.tune_walk(geom, D, wf, 20000, probe_walkers=8192, probe_samples=16)
    assert set(ms) == {4, 16, 32} and best in ms and all(v > 0 for v in ms.values())
    pw = gpu.PlannedWalk(geom, D, wf, 4096)
    assert pw.meas_block == best                    # the cached choice is used
    a = pw.run(seed=5)["signal"]
    b = gpu.walk(geom, D, wf, 4096, seed=5, meas_block=4)["signal"]
    assert np.max(np.abs(a - b)) <= 1e-5            # any block walks the same trajectories
    pw.close()
    autotune.clear()


# ---------------------------------------------------------------------------
# Closed two-compartment cells: Monte-Carlo exchange time against the exact
# finite-diffusion eigenvalue, and the equilibrium partition (detailed balance).
# Slab: lambda tan(lambda L/2) = 2 kappa / D.  Shells: the lowest root of the
# 3x3 determinant of {reflecting R_out, flux continuity at R_in, permeable jump
# at R_in} for the radial (Bessel / spherical Bessel) modes; tau = 1/(D lambda^2).
# Same cells, kappa, walker count and 2 % tolerances as the reference's ladder
# (tests/validation/test_permeability_ladder.py:26-49).
# ---------------------------------------------------------------------------
def _lowest_root(f, hi, n=4000):
    from scipy.optimize import brentq
    xs = np.linspace(hi * 1e-4, hi, n)
    vs = [f(x) for x in xs]
    for a, b, fa, fb in zip(xs[:-1], xs[1:], vs[:-1], vs[1:]):
        if np.sign(fa) != np.sign(fb):
            return brentq(f, a, b)
    raise RuntimeError("no root")


def _tau_slab(L, kappa):
    lam = _lowest_root(lambda l: l * np.tan(l * L / 2) - 2 * kappa / D, np.pi / L * 0.999)
    return 1.0 / (D * lam ** 2)


def _tau_shell(Rin, Rout, kappa, kind):
    from scipy.special import jv, yv, spherical_jn, spherical_yn

    def det(l):
        a, b = l * Rin, l * Rout
        if kind == "sphere":
            F0, F0p = spherical_jn(0, a), spherical_jn(0, a, True)
            G0, G0p = spherical_yn(0, a), spherical_yn(0, a, True)
            Fbp, Gbp = spherical_jn(0, b, True), spherical_yn(0, b, True)
        else:
            F0, F0p, G0, G0p = jv(0, a), -jv(1, a), yv(0, a), -yv(1, a)
            Fbp, Gbp = -jv(1, b), -yv(1, b)
        M = np.array([[0.0, Fbp, Gbp],
                      [F0p, -F0p, -G0p],
                      [-D * l * F0p - kappa * F0, kappa * F0, kappa * G0]])
        return np.linalg.det(M)

    return 1.0 / (D * _lowest_root(det, 1.2e5) ** 2)


def _f_inside(gpu, geom, t, dt, n):
    n_t = max(2, int(round(t / dt)))
    wf = gpu.Waveform(G=np.zeros((1, n_t, 3), np.float32), dt=float(dt), echo_idx=n_t - 1)
    _, _, cf = gpu.simulate(n, D, wf, geom, seed=3, return_compartments="final")
    return float((np.asarray(cf) == 0).mean()), n_t * dt


def _exchange(gpu, geom, tau_ex, feq, R, n):
    dt = (R / 50) ** 2 / (6 * D)
    ts, fs = [], []
    for m in (0.6, 0.9, 1.2, 1.5):
        f, ta = _f_inside(gpu, geom, m * tau_ex, dt, n)
        ts.append(ta)
        fs.append(f)
    y = (np.array(fs) - feq) / (1 - feq)
    tau = -1.0 / np.polyfit(np.array(ts), np.log(y), 1)[0]
    f_eq, _ = _f_inside(gpu, geom, 8 * tau_ex, (0.1 * R) ** 2 / (6 * D), n)
    return tau, f_eq


def test_permeable_slab_exchange_time(gpu):
    L, kappa = 20e-6, 20e-6
    tau_ex = _tau_slab(L, kappa)
    tau, f_eq = _exchange(gpu, gpu.PermeableSlab1D(L, kappa), tau_ex, 0.5, L / 2, 500_000)
    print(f"slab: tau {tau * 1e3:.2f} ms vs exact {tau_ex * 1e3:.2f} ms; f_eq {f_eq:.4f}")
    assert abs(tau - tau_ex) / tau_ex < 0.02
    assert abs(f_eq - 0.5) < 0.02


@pytest.mark.parametrize("kind,feq", [("cylinder", 0.25), ("sphere", 0.125)])
def test_permeable_shell_exchange_time(gpu, kind, feq):
    R, kappa = 10e-6, 20e-6
    tau_ex = _tau_shell(R, 2 * R, kappa, kind)
    tau, f_eq = _exchange(gpu, gpu.PermeableShell(R, 2 * R, kappa, kind=kind), tau_ex, feq, R,
                          500_000)
    print(f"{kind} shell: tau {tau * 1e3:.2f} ms vs exact {tau_ex * 1e3:.2f} ms; f_eq {f_eq:.4f}")
    assert abs(tau - tau_ex) / tau_ex < 0.02
    assert abs(f_eq - feq) < 0.02


def test_ellipsoid_sphere_limit_and_anisotropy(gpu):
    """a = b = c reproduces Sphere(a) within Monte-Carlo error; a prolate
    ellipsoid attenuates more along its long axis than across it."""
    wf = ds.set_b(ds.pgse(0.01, 0.03, 0.05, np.eye(3), 200, slew_rate=np.inf), [2e9] * 3)
    n = 200_000
    Se = gpu.simulate(n, D, wf, gpu.Ellipsoid((4e-6, 4e-6, 4e-6)), seed=1)
    Ss = gpu.simulate(n, D, wf, gpu.Sphere(4e-6), seed=2)
    assert np.all(np.abs(Se - Ss) <= 4 * np.sqrt(2 * 0.5 / n))
    Sp = gpu.simulate(n, D, wf, gpu.Ellipsoid((2e-6, 2e-6, 8e-6)), seed=3)
    assert Sp[2] < Sp[0] - 0.05 and abs(Sp[0] - Sp[1]) < 4 * np.sqrt(2 * 0.5 / n)


def test_free_gpu_memory_closes_live_objects(gpu):
    """free_gpu_memory(aggressive=True) closes every live planned walk, batched
    walk, network and trajectory set (dmipy_sim/gpu.py:93-134's aggressive mode);
    a closed object refuses further use."""
    wf = small_pgse(n_t=20)
    pw = gpu.PlannedWalk(ds.Sphere(3e-6), D, wf, 256)
    bw = gpu.BatchedWalk([dict(geometry=ds.Sphere(3e-6), diffusivity=D, n_walkers=128)], wf)
    pw.run(seed=1)
    assert gpu.free_gpu_memory() == 0              # both still referenced: nothing closed
    assert gpu.free_gpu_memory(aggressive=True) >= 2
    assert pw._h is None and bw._h is None
    with pytest.raises((ValueError, RuntimeError, AttributeError)):
        pw.run(seed=2)
