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
