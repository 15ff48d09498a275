"""Pins of the replay oracle (oracle/replay_oracle.py) by closed forms, and the
product's producer constants against the oracle's restatement."""
import math

import numpy as np
import pytest

from oracle import replay_oracle as ro

import dmipy_sim_amd as ds
from dmipy_sim_amd import trajectories as tj

GAMMA = 267.513e6


def _pgse(n_t=200, T=0.04, b=1e9, dirs=((1, 0, 0), (0, 1, 0), (0, 0.6, 0.8))):
    wf = ds.pgse(0.01, 0.02, 0.05, np.asarray(dirs, float), n_t, slew_rate=np.inf)
    return ds.set_b(wf, [b] * len(dirs))


def test_static_walkers_under_zero_moment_waveform():
    wf = _pgse()
    G = np.asarray(wf.G, np.float64)
    assert np.allclose(G.sum(axis=1), 0.0, atol=1e-12)      # square lobes: zero net moment
    rng = np.random.default_rng(0)
    c = rng.normal(0, 5e-6, (50, 3))
    traj = np.repeat(c[:, None, :], G.shape[1], axis=1)
    S, phi, _ = ro.replay(traj, wf.dt, wf.G, wf.dt)
    assert np.max(np.abs(phi)) < 1e-9
    np.testing.assert_allclose(S, 1.0, atol=1e-12)


def test_uniform_motion_phase_by_summation_by_parts():
    """r_w(t_k) = v_w k dt: phi = GAMMA dt^2 v . [(n-1) Q_{n-1} - sum_{k<n-1} Q_k],
    Q_k = sum_{j<=k} G_j (Abel summation, no GEMM)."""
    wf = _pgse(n_t=150)
    G = np.asarray(wf.G, np.float64)
    n = G.shape[1]
    rng = np.random.default_rng(1)
    v = rng.normal(0, 1e-4, (40, 3))
    k = np.arange(n, dtype=np.float64)
    traj = v[:, None, :] * (k * wf.dt)[None, :, None]
    _, phi, _ = ro.replay(traj, wf.dt, wf.G, wf.dt)
    Q = np.cumsum(G, axis=1)
    A = (n - 1) * Q[:, -1, :] - Q[:, :-1, :].sum(axis=1)        # (n_meas, 3)
    ref = GAMMA * wf.dt ** 2 * (A @ v.T)
    np.testing.assert_allclose(phi, ref, rtol=1e-9, atol=1e-9 * np.max(np.abs(ref)))


def test_relaxation_weights_closed_form():
    n_t, dt = 120, 1e-4
    G = np.zeros((2, n_t, 3), np.float32)
    chi = np.ones(n_t)
    chi[40:70] = 0.0
    traj = np.zeros((8, n_t, 3))
    S, _, lw = ro.replay(traj, dt, G, dt, chi=chi, T2=0.05, T1=0.8)
    ref = math.exp(-(dt / 0.05) * 90 - (dt / 0.8) * 30)
    np.testing.assert_allclose(S, ref, rtol=1e-12)
    # boundary log-weight: constant d_w per step with chi = 1 -> exp((rho/D) d_w n_t)
    d = -np.linspace(1e-9, 5e-8, 8)
    dlog = np.repeat(d[:, None], n_t, axis=1)
    S2, _, lw2 = ro.replay(traj, dt, G, dt, dlog=dlog, rho=2e-5, D=2e-9)
    np.testing.assert_allclose(S2, np.mean(np.exp((2e-5 / 2e-9) * d * n_t)), rtol=1e-12)
    assert lw2.shape == (2, 8)


def test_resampling_linear_waveform_and_nearest_chi():
    dt_wf, n_wf = 2e-4, 51
    t_wf = np.arange(n_wf) * dt_wf
    G = np.zeros((1, n_wf, 3), np.float32)
    G[0, :, 1] = (3.0 * t_wf).astype(np.float32)               # linear in time
    chi = (np.arange(n_wf) % 10 < 5).astype(float)
    dt_tr, n_tr = 1e-4, 120                                     # finer grid, runs past the end
    Gt, chi_r = ro.resample(G, dt_wf, dt_tr, n_tr, chi)
    t_tr = np.arange(n_tr) * dt_tr
    inside = t_tr <= t_wf[-1] + 1e-15
    np.testing.assert_allclose(Gt[0, inside, 1], 3.0 * t_tr[inside], rtol=1e-6, atol=1e-9)
    assert np.all(Gt[0, ~inside, :] == 0)
    idx = np.round(t_tr / dt_wf).astype(int)
    exp_chi = np.where(idx < n_wf, chi[np.clip(idx, 0, n_wf - 1)], 0.0)
    np.testing.assert_array_equal(chi_r[0], exp_chi)


@pytest.mark.parametrize("geom,perm,R", [
    (ds.FreeDiffusion(), False, None),
    (ds.Box1D(8e-6), False, 8e-6),
    (ds.Cylinder(3e-6, (0, 0, 1), surface_relaxivity_t2=1e-5), False, 3e-6),
    (ds.Sphere(4e-6, permeability=2e-5), True, 4e-6),
])
@pytest.mark.parametrize("save_relax", [False, True])
def test_producer_constants_match(geom, perm, R, save_relax):
    D, T_max, dt = 2e-9, 0.04, 0.04 / 199
    o = ro.producer_constants(R, perm, D, T_max, dt, save_relax)
    p = tj.producer_plan(geom, D, T_max, dt, save_relax)
    assert p["n_t"] == o["n_t"] == 200
    assert p["sub_steps"] == o["sub_steps"]
    assert p["step_len"] == o["step_len"]
    assert p["dt_actual"] == o["dt"]
    if R is not None:
        # the rule's definition: the step is at most R/6 (R/25 when permeable)
        frac = 25.0 if perm else 6.0
        assert math.sqrt(6 * D * o["dt_sim"]) <= R / frac * (1 + 1e-12)
        assert o["sub_steps"] == 1 or math.sqrt(6 * D * o["dt"] / (o["sub_steps"] - 1)) > R / frac
    if isinstance(geom, ds.FreeDiffusion):
        assert p["wall"] == 0 and p["rho_over_D"] == 0.0
    elif save_relax:
        assert p["rho_over_D"] == 1.0
