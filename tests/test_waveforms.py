"""Waveform builders against closed forms and the reference's integer timing rules
(tests/test_general_waveform.py:90-106, tests/test_free_1d.py:14-35,
tests/test_ogse_waveform.py:28-71, README.md:37)."""
import numpy as np
import pytest

import dmipy_sim_amd as ds

GAMMA = ds.GAMMA


def test_pgse_timing_integers_readme_workload():
    """pgse(0.01, 0.04, 0.05, n_t=300): dt = 0.05/299, n_pulse = 60, n_DELTA = 239,
    slew 200 T/m/s -> n_rise = round(0.05/200/dt) = 1."""
    wf = ds.pgse(0.01, 0.04, 0.05, [[1, 0, 0]], 300)
    dt = 0.05 / 299
    assert wf.dt == pytest.approx(dt, rel=1e-15)
    g = wf.G[0, :, 0]
    nz = np.nonzero(g)[0]
    assert nz[0] == 1 and g[0] == 0.0          # one-sample ramp starts at 0
    first = nz[nz < 150]
    second = nz[nz >= 150]
    assert first[-1] == 59 and second[0] == 240 and second[-1] == 298
    assert np.all(g[second] <= 0) and np.all(g[first] >= 0)
    assert g[30] == np.float32(0.05)


def test_calc_b_matches_square_pgse_closed_form():
    wf = ds.pgse(10e-3, 40e-3, 0.05, [[1., 0., 0.]], 2000, slew_rate=np.inf)
    b = ds.calc_b(wf)[0]
    assert b == pytest.approx((GAMMA * 0.05 * 10e-3) ** 2 * (40e-3 - 10e-3 / 3), rel=0.01)


def test_set_b_round_trip():
    b = np.linspace(1, 2e9, 100)
    wf = ds.set_b(ds.pgse(8e-3, 50e-3, 1.0, np.tile([1.0, 0, 0], (100, 1)), 1000), b)
    np.testing.assert_allclose(ds.calc_b(wf), b, rtol=1e-3)


def test_set_b_warns_on_mm2_units():
    wf = ds.pgse(8e-3, 50e-3, 1.0, [[1.0, 0, 0]], 100)
    with pytest.warns(UserWarning, match="s/mm"):
        ds.set_b(wf, [1000.0])


def _q(wf):
    G = np.asarray(wf.G)[0, :, 0].astype(np.float64)
    return GAMMA * np.concatenate([[0.0], np.cumsum(0.5 * (G[1:] + G[:-1])) * wf.dt])


@pytest.mark.parametrize("freq", [25.0, 50.0, 100.0])
def test_cosine_ogse_refocused_and_dc_free(freq):
    wf = ds.ogse(freq, 0.080, 0.05, [[1, 0, 0]], 1601, kind='cosine')
    q = _q(wf)
    qmax = np.max(np.abs(q)) + 1e-30
    assert abs(q[-1]) / qmax < 1e-2
    assert abs(np.mean(q)) / qmax < 0.05


def test_ogse_snapping_warns():
    with pytest.warns(UserWarning, match="snapped"):
        ds.ogse(37.0, 0.080, 0.05, [[1, 0, 0]], 401)


def test_cpmg_echo_indices_and_sign_flips():
    wf = ds.cpmg(4, 10e-3, 0.02, [[1, 0, 0]], n_t_per_echo=50)
    assert wf.G.shape == (1, 201, 3)
    np.testing.assert_array_equal(wf.echo_indices, [50, 100, 150, 200])
    g = wf.G[0, :, 0]
    t = np.arange(201) * wf.dt
    expect = np.where(np.floor((t / 10e-3) + 0.5) % 2 == 0, 0.02, -0.02).astype(np.float32)
    np.testing.assert_array_equal(g, expect)


def test_rotate_and_tile():
    wf = ds.pgse(0.01, 0.04, 0.05, [[1, 0, 0]], 50, slew_rate=np.inf)
    r = ds.rotate_waveform(wf, theta=np.pi / 2)
    np.testing.assert_allclose(r.G[0, :, 2], -wf.G[0, :, 0], atol=1e-9)
    t = ds.tile_waveform(wf, 3)
    assert t.G.shape == (3, 50, 3)


# ---------------------------------------------------------------- stimulated echo, b-tensors
def test_pgste_timing_mask_and_b():
    """PGSTE (waveforms.py:166-262): lobes at 0 and n_pulse + round(TM/dt), the
    mixing time stored (chi_perp 0), the 0.5 stimulated-echo flag, refocused
    (sum of G = 0), and for square lobes b = (gamma G delta)^2 (Delta - delta/3)
    with Delta = delta + TM (the lobe spacing) within 1 %."""
    delta, TM, n_t = 5e-3, 40e-3, 1000
    wf = ds.pgste(delta, TM, 0.05, [[1, 0, 0]], n_t, slew_rate=np.inf)
    dt = (2 * delta + TM) / (n_t - 1)
    n_pulse = max(1, round(delta / dt))
    i_recall = n_pulse + round(TM / dt)
    assert wf.stimulated_echo and wf.TM == TM and wf.dt == pytest.approx(dt)
    chi = np.asarray(wf.chi_perp)
    assert chi[:n_pulse].all() and not chi[n_pulse:i_recall].any() and chi[i_recall:].all()
    G = np.asarray(wf.G)
    assert np.all(G[0, :n_pulse, 0] == np.float32(0.05)) and np.all(G[0, n_pulse:i_recall] == 0)
    assert np.all(G[0, i_recall:i_recall + n_pulse, 0] == np.float32(-0.05))
    assert abs(G.sum(axis=1)[0, 0]) < 1e-5
    b = ds.calc_b(wf)[0]
    d_eff, D_eff = n_pulse * dt, i_recall * dt
    assert b == pytest.approx((ds.GAMMA * 0.05 * d_eff) ** 2 * (D_eff - d_eff / 3), rel=1e-2)
    wf2 = ds.set_b(wf, [1e9])
    assert wf2.stimulated_echo and np.array_equal(wf2.chi_perp, wf.chi_perp)


@pytest.mark.parametrize("N", [1, 2, 3, 4])
def test_trapezoidal_ogse_matches_closed_form_b(N):
    """The sampled trapezoidal N-lobe OGSE against Eq. 2 of Drobnjak et al. 2016
    (b_trapezoidal_ogse), with the ramp and lobe times the samples realise; odd N
    carries the net-moment term, even N refocuses within each block."""
    delta, DELTA, G, n_t, tr = 12e-3, 32e-3, 0.08, 6000, 0.4e-3
    wf = ds.trapezoidal_ogse(N, delta, DELTA, G, [[0, 1, 0]], n_t, rise_time=tr)
    dt = wf.dt
    n_lobe = max(1, round(delta / (N * dt)))
    n_rise = min(round(tr / dt), n_lobe // 2)
    b_num = ds.calc_b(wf)[0]
    b_ref = ds.b_trapezoidal_ogse(N, n_lobe * N * dt, round(DELTA / dt) * dt, G, n_rise * dt)
    assert b_num == pytest.approx(b_ref, rel=1e-2)
    Gs = np.asarray(wf.G)[0, :, 1]
    assert abs(Gs.sum()) < 1e-4 and np.max(np.abs(Gs)) == pytest.approx(G, rel=1e-6)
    # square limit: the closed form with zero ramp
    wf0 = ds.trapezoidal_ogse(N, delta, DELTA, G, [[0, 1, 0]], n_t, rise_time=0.0)
    assert ds.calc_b(wf0)[0] == pytest.approx(
        ds.b_trapezoidal_ogse(N, n_lobe * N * dt, round(DELTA / dt) * dt, G, 0.0), rel=1e-2)


def test_btensor_trace_and_linear_encoding():
    """trace(B) = calc_b (waveforms.py:519-543); a PGSE along any direction is a
    linear encoding: b_delta = 1, b_eta = 0, B = b n n^T."""
    for bv in ([1, 0, 0], [0, 0, 1], [0.6, 0.8, 0.0]):
        wf = ds.pgse(10e-3, 40e-3, 0.06, [bv], 1000)
        B = ds.calc_btensor(wf)
        assert np.trace(B[0]) == pytest.approx(ds.calc_b(wf)[0], rel=1e-5)
        b, bd, be = ds.btensor_invariants(B)
        assert abs(bd[0] - 1) < 1e-3 and abs(be[0]) < 1e-3
        n = np.asarray(bv, np.float64) / np.linalg.norm(bv)
        np.testing.assert_allclose(B[0], b[0] * np.outer(n, n), atol=2e-5 * b[0])
    wf = ds.trapezoidal_ogse(2, 10e-3, 40e-3, 0.06, [[1, 0, 0]], 1000)
    assert np.trace(ds.calc_btensor(wf)[0]) == pytest.approx(ds.calc_b(wf)[0], rel=1e-5)


def test_ste_and_pte_tensor_shapes():
    """STE: B = (b/3) I, b_delta = 0; PTE: eigenvalues (b/2, b/2, 0) with the zero
    along the plane normal, b_delta = -1/2; both keep their shape under set_b
    (waveforms.py:598-720)."""
    wf = ds.ste(10e-3, 40e-3, 0.06, 1000)
    B = ds.calc_btensor(wf)[0]
    b = np.trace(B)
    np.testing.assert_allclose(np.diag(B), b / 3, rtol=1e-5)
    np.testing.assert_allclose(B - np.diag(np.diag(B)), 0.0, atol=1e-5 * b)
    _, bd, be = ds.btensor_invariants(B[None])
    assert abs(bd[0]) < 1e-3 and abs(be[0]) < 1e-3
    for normal in ([0, 0, 1], [1, 0, 0], [1, 1, 1]):
        n = np.asarray(normal, np.float64) / np.linalg.norm(normal)
        Bp = ds.calc_btensor(ds.pte(10e-3, 40e-3, 0.06, n, 1000))[0]
        w, V = np.linalg.eigh(Bp)
        bp = np.trace(Bp)
        np.testing.assert_allclose(w[1:], bp / 2, rtol=1e-5)
        assert abs(w[0]) < 1e-6 * bp and abs(abs(V[:, 0] @ n) - 1) < 1e-6
        _, bd, _ = ds.btensor_invariants(Bp[None])
        assert abs(bd[0] + 0.5) < 1e-3
    for wf0 in (ds.ste(10e-3, 40e-3, 0.06, 1000), ds.pte(10e-3, 40e-3, 0.06, [0, 0, 1], 1000)):
        b_out, bd, _ = ds.btensor_invariants(ds.calc_btensor(ds.set_b(wf0, 1e9)))
        assert b_out[0] == pytest.approx(1e9, rel=1e-4)
    assert ds.btensor_invariants(np.zeros((1, 3, 3)))[1][0] == 0.0
