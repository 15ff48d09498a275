"""The NumPy array-at-a-time restatement (the CPU baseline's second variant,
oracle/numpy_walk.py) walks the same physics as the C oracle: on a permeable
packing, with independent random streams, its signal, displacement spread and
intra-cylinder fraction agree with the oracle's to Monte-Carlo precision."""
import numpy as np

import dmipy_sim_amd as ds
from substrates import D, packed_cyl


def test_numpy_walk_matches_the_oracle_statistically(oracle):
    from oracle import numpy_walk
    geom = packed_cyl(kappa=5e-4, n=8, vf=0.45)
    n_t, dt = 4, 1e-3
    G = np.zeros((3, n_t, 3), np.float32)
    G[0, :, 0] = 0.3
    G[1, :, 1] = -0.25
    G[2, :2, 0], G[2, 2:, 0] = 0.4, -0.4
    wf = ds.Waveform(G=G, dt=dt, echo_idx=n_t - 1)
    p = oracle.oracle_plan(geom, D, dt)
    N = 3000
    r0 = np.array(geom.init_positions(N, 3), np.float32)
    o = oracle.run_oracle(geom, D, wf, N, seed=5, r0=r0, cell_form=0)
    sig, r = numpy_walk.packed_cylinder_walk(
        geom._centers_np, geom._radii_np, p["cell"], p["kappa_over_D"], p["eps"], p["nudge"],
        p["step_len"], p["gamma_dt_sub"], G, p["n_sub"], r0, seed=7)
    assert np.abs(o["signal"]).min() < 0.5          # the phase is not trivial
    se = np.sqrt(0.5 / N) * 2
    assert np.all(np.abs(sig - o["signal"]) < 4 * se), (sig, o["signal"])
    msd_o = np.mean(np.sum((o["final_pos"][:, :2] - r0[:, :2]) ** 2, 1))
    msd_n = np.mean(np.sum((r[:, :2] - r0[:, :2]) ** 2, 1))
    assert abs(msd_n / msd_o - 1) < 0.08, (msd_n, msd_o)
    L = p["cell"]

    def intra(x):
        q = x[:, None, :2] - geom._centers_np[None]
        q -= L * np.floor(q / L + 0.5)
        return np.mean(np.any(np.sum(q * q, 2) < geom._radii_np ** 2, 1))
    fi_o, fi_n = intra(o["final_pos"]), intra(r)
    assert abs(fi_o - fi_n) < 4 * np.sqrt(fi_o * (1 - fi_o) / N) + 1e-3, (fi_o, fi_n)
    assert fi_o > 0.01                               # walkers did cross into the cylinders
