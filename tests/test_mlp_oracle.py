"""Pins of the surrogate MLP's CPU restatement (oracle/mlp_oracle.py).

- bfloat16 rounding against torch's own float32 -> bfloat16 conversion and
  hand-worked ties;
- the exact (unrounded) backward pass against central finite differences of
  the forward pass, so a dropped GELU' factor, a transposed weight or a wrong
  loss scale fails;
- the Adam step against torch.optim.Adam;
- the waveform embedding against Parseval's identity (the bins sum to the
  rectangle-rule b-value) and against calc_b.
"""
import numpy as np
import pytest
import torch

from oracle import mlp_oracle as mo

import dmipy_sim_amd as ds
from dmipy_sim_amd import surrogate as sg


def test_bf16_round_matches_torch_and_ties():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(10000) * 10 ** rng.uniform(-8, 8, 10000),
                        [0.0, -0.0, 1.0, -2.5, 3.0e38, 1e-40]]).astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy().astype(np.float64)
    np.testing.assert_array_equal(mo.bf16_round(x), ref)
    # 1 + 2^-8 is halfway between 1 and 1 + 2^-7: ties go to the even mantissa.
    assert mo.bf16_round(np.float32(1 + 2.0 ** -8)) == 1.0
    assert mo.bf16_round(np.float32(1 + 3 * 2.0 ** -8)) == 1 + 2.0 ** -6


def test_gelu_grad_is_derivative():
    x = np.linspace(-6, 6, 1001)
    h = 1e-6
    fd = (mo.gelu(x + h) - mo.gelu(x - h)) / (2 * h)
    np.testing.assert_allclose(mo.gelu_grad(x), fd, atol=1e-8)
    # tanh-form GELU against the erf form: within 1e-3 everywhere
    from scipy.special import erf
    np.testing.assert_allclose(mo.gelu(x), 0.5 * x * (1 + erf(x / np.sqrt(2))), atol=1e-3)


@pytest.mark.parametrize("n_in,width,depth", [(5, 8, 1), (7, 6, 3)])
def test_exact_gradient_matches_finite_differences(n_in, width, depth):
    rng = np.random.default_rng(n_in)
    n = mo.num_params(n_in, width, depth)
    p = rng.standard_normal(n) * 0.5
    x = rng.standard_normal((9, n_in))
    t = rng.standard_normal(9)
    g, loss = mo.gradients(p, x, t, n_in, width, depth, bf16=False)
    y = mo.forward(p, x, n_in, width, depth, bf16=False)
    assert loss == pytest.approx(np.mean((y - t) ** 2))
    h = 1e-6
    fd = np.empty(n)
    for i in range(n):
        e = np.zeros(n)
        e[i] = h
        lp = np.mean((mo.forward(p + e, x, n_in, width, depth, bf16=False) - t) ** 2)
        lm = np.mean((mo.forward(p - e, x, n_in, width, depth, bf16=False) - t) ** 2)
        fd[i] = (lp - lm) / (2 * h)
    np.testing.assert_allclose(g, fd, rtol=1e-5, atol=1e-7)


def test_layout_matches_torch_module():
    """Public layout and forward pass against a torch.nn stack (float64, no rounding)."""
    n_in, width, depth = 6, 8, 3
    rng = np.random.default_rng(5)
    p = rng.standard_normal(mo.num_params(n_in, width, depth))
    Ws, bs, w_out, b_out = mo.unpack(p, n_in, width, depth)
    layers = []
    for l in range(depth):
        lin = torch.nn.Linear(Ws[l].shape[0], width).double()
        lin.weight.data = torch.from_numpy(Ws[l].T.copy())
        lin.bias.data = torch.from_numpy(bs[l].copy())
        layers += [lin, torch.nn.GELU(approximate="tanh")]
    out = torch.nn.Linear(width, 1).double()
    out.weight.data = torch.from_numpy(w_out[None, :].copy())
    out.bias.data = torch.tensor([b_out], dtype=torch.float64)
    net = torch.nn.Sequential(*layers, out)
    x = rng.standard_normal((11, n_in))
    ref = net(torch.from_numpy(x)).detach().numpy()[:, 0]
    np.testing.assert_allclose(mo.forward(p, x, n_in, width, depth, bf16=False), ref, rtol=1e-12)
    np.testing.assert_allclose(mo.pack(Ws, bs, w_out, b_out), p)


def test_adam_matches_torch():
    rng = np.random.default_rng(2)
    p0 = rng.standard_normal(50)
    grads = [rng.standard_normal(50) for _ in range(5)]
    tp = torch.nn.Parameter(torch.from_numpy(p0.copy()))
    opt = torch.optim.Adam([tp], lr=1e-2, betas=(0.9, 0.999), eps=1e-8)
    p, m, v = p0.copy(), np.zeros(50), np.zeros(50)
    for k, g in enumerate(grads, start=1):
        tp.grad = torch.from_numpy(g.copy())
        opt.step()
        p, m, v = mo.adam_step(p, g, m, v, k, 1e-2)
    np.testing.assert_allclose(p, tp.detach().numpy(), rtol=1e-12, atol=1e-14)


def test_waveform_embedding_parseval_and_calc_b():
    rng = np.random.default_rng(1)
    wf = sg.random_waveforms(rng, n_meas=8, n_t=250)
    f = sg.waveform_features(wf)
    assert f.shape == (8, 64)
    q = ds.GAMMA * np.cumsum(np.asarray(wf.G, np.float64), axis=1) * wf.dt
    b_rect = np.sum(q ** 2, axis=(1, 2)) * wf.dt * 1e-9
    b_perp = np.sum(q[:, :, :2] ** 2, axis=(1, 2)) * wf.dt * 1e-9
    # Parseval: the transverse bins sum to the transverse b-value, x and y split it
    np.testing.assert_allclose(f[:, :56].sum(axis=1), b_perp, rtol=1e-10)
    np.testing.assert_allclose(f[:, 56], b_perp, rtol=1e-10)
    np.testing.assert_allclose(f[:, 56] + f[:, 57], b_rect, rtol=1e-10)
    np.testing.assert_allclose(f[:, 60] + f[:, 61], f[:, 56], rtol=1e-10)
    assert not np.any(f[0])                       # the b = 0 row
    b = np.asarray(ds.calc_b(wf), np.float64) * 1e-9
    np.testing.assert_allclose(f[1:, 56] + f[1:, 57], b[1:], rtol=2e-2)
    # 1 <= <q^4>/<q^2>^2 <= max q^2 / <q^2> (Cauchy-Schwarz, q^4 <= max q^2 . q^2)
    assert np.all(f[1:, 58] >= 1.0) and np.all(f[1:, 58] <= f[1:, 59] * (1 + 1e-12))
    # a z-only waveform has no transverse part
    G = np.zeros_like(wf.G)
    G[:, :, 2] = wf.G[:, :, 0]
    fz = sg.waveform_features(ds.Waveform(G=G, dt=wf.dt, echo_idx=wf.echo_idx))
    assert not np.any(fz[:, :57]) and not np.any(fz[:, 58:])


def test_waveform_embedding_closed_form_of_one_sine():
    """G_x = A sin(2 pi k t / T) over whole periods: q = c (1 - cos), so
    <q^4>/<q^2>^2 = (35/8) / (3/2)^2, max q^2 / <q^2> = 4 / (3/2), and the
    spectrum holds b in the ratio 2:1 at frequency 0 and k: b-weighted mean
    frequency k/3, root-mean-square frequency k/sqrt(3)."""
    n_t, T, k = 4000, 0.04, 7
    dt = T / n_t
    t = (np.arange(n_t) + 1) * dt
    G = np.zeros((1, n_t, 3))
    G[0, :, 0] = 0.02 * np.sin(2 * np.pi * k * t / T)
    f = sg.waveform_features(ds.Waveform(G=G, dt=dt, echo_idx=n_t - 1))[0]
    assert f[58] == pytest.approx((35 / 8) / 2.25, rel=2e-3)
    assert f[59] == pytest.approx(4 / 1.5, rel=2e-3)
    assert f[62] * 40 == pytest.approx(k / 3, rel=2e-3)
    assert f[63] * 40 == pytest.approx(k / np.sqrt(3), rel=2e-3)
    assert f[61] == 0.0 and f[57] == 0.0


def test_features_shape_and_prior():
    rng = np.random.default_rng(3)
    p = sg.sample_params(rng)
    for k, (lo, hi) in sg.PRIOR.items():
        assert lo <= p[k] <= hi or (k == "rho" and p[k] == 0.0)
    rhos = [sg.sample_params(rng)["rho"] for _ in range(400)]
    assert 0.15 < np.mean(np.equal(rhos, 0.0)) < 0.35          # RHO_ZERO of the instances
    x = sg.features(p, sg.random_waveforms(rng, n_meas=5))
    assert x.shape == (5, sg.N_IN) and x.dtype == np.float32
    assert np.all(np.abs(x[:, :6]) <= 1.5)
    geom = sg.build_substrate(p, seed=4)
    assert np.min(geom._radii_np) >= sg.R_MIN * (1 - 1e-6)


def test_generator_matches_the_packed_workload_family():
    """The training generator draws the packed-cylinder workload's family: 64
    cylinders, 90 free waveforms of n_t = 2000 over 25 ms with zero net moment
    and max |G| in [0.05, 0.3] T/m; a volume-fraction prior containing 0.6."""
    rng = np.random.default_rng(8)
    wf = sg.random_waveforms(rng)
    G = np.asarray(wf.G, np.float64)
    assert G.shape == (90, 2000, 3)
    assert wf.dt == pytest.approx(0.025 / 1999)
    np.testing.assert_allclose(G.sum(axis=1), 0.0, atol=1e-4)
    gmax = np.linalg.norm(G, axis=2).max(axis=1)
    assert gmax[0] == 0.0 and np.all((gmax[1:] >= 0.05 * 0.999) & (gmax[1:] <= 0.3 * 1.001))
    assert sg.PRIOR["vf"][0] <= 0.6 <= sg.PRIOR["vf"][1]
    geom = sg.build_substrate(sg.sample_params(rng), seed=2)
    assert len(geom._radii_np) == 64


@pytest.mark.parametrize("n_meas,max_meas", [(1, 64), (45, 64), (64, 64), (65, 64), (90, 64), (130, 64),
                                             (90, 128), (128, 128), (130, 128), (300, 128)])
def test_split_measurements_covers_every_measurement(n_meas, max_meas):
    """Batched instances take at most max_meas measurements (64, or 128 for the
    permeable packings the generator walks): a waveform is walked as equal-size
    sets, the last padded with zero-gradient rows, which the concatenation drops
    again."""
    rng = np.random.default_rng(n_meas)
    wf = sg.random_waveforms(rng, n_meas=n_meas, n_t=40)
    parts = sg.split_measurements(wf, max_meas)
    k = parts[0].G.shape[0]
    assert k <= max_meas and all(p.G.shape == (k, 40, 3) for p in parts)
    assert len(parts) == -(-n_meas // max_meas)
    cat = np.concatenate([p.G for p in parts])
    np.testing.assert_array_equal(cat[:n_meas], wf.G)
    assert not np.any(cat[n_meas:])
    assert all(p.dt == wf.dt for p in parts)


def test_instance_params_without_relaxivity():
    """The packed workload's substrate (no surface relaxivity) is an input like
    any other: rho = 0, floored in the relaxivity feature."""
    from dmipy_sim_amd import configs
    g = configs.packed_cylinders_substrate()
    p = sg.instance_params(g, 2e-9)
    assert p["rho"] == 0.0 and p["kappa"] == 1e-5
    assert abs(p["vf"] - 0.6) < 1e-9 and abs(p["mean_r"] - np.mean(g._radii_np)) < 1e-15
    x = sg.substrate_features(p)
    assert np.all(np.isfinite(x)) and np.all(np.abs(x) <= 2.0)


def test_generator_inputs_do_not_depend_on_the_worker_pool(monkeypatch):
    """Every instance is built from its own seeds drawn in order, so the inputs
    (and substrates) are the same with the spawned worker pool and inline; the
    walk is replaced by a host stub that records what it is given."""
    seen = []

    def fake_walk(instances, waveforms, n_walkers, seed):
        seen.append([np.asarray(i["geometry"]._radii_np) for i in instances])
        return np.zeros((len(instances), waveforms[0].G.shape[0])), {"total_ms": 0.0, "walker_steps": 0}

    monkeypatch.setattr(sg, "walk_signals", fake_walk)
    Xp, _, ip = sg.generate(12, 64, seed=5, n_meas=6, n_t=60, chunk=5, workers=3)
    radii_pool = [r for c in seen for r in c]
    seen.clear()
    Xi, _, ii = sg.generate(12, 64, seed=5, n_meas=6, n_t=60, chunk=12, workers=1)
    np.testing.assert_array_equal(Xp, Xi)
    assert all(np.array_equal(a, b) for a, b in zip(radii_pool, seen[0]))
    assert len(ip["instances"]) == 12 and ip["params"][3] == ii["params"][3]
    _, _, ik = sg.generate(3, 64, seed=5, n_meas=6, n_t=60, keep_instances=False, workers=1)
    assert ik["instances"] == [] and len(ik["params"]) == 3
