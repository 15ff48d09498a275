"""Randomised parity: many small walks with random substrate sizes, diffusivities,
time steps, relaxivities and permeabilities, GPU against the oracle on the same
Philox stream.  Spreads the operands of the wall arithmetic over several decades
(positions, step lengths, radii from 0.3 to 30 um; steps from R/40 to R) so the
exact-arithmetic routines are exercised far beyond the fixed parity cases; half the
trials run the phase product on the matrix cores (float16 hi/lo split operands)."""
import numpy as np
import pytest

import dmipy_sim_amd as ds

pytestmark = pytest.mark.gpu


def _random_case(rng, kind):
    R = float(10 ** rng.uniform(-6.5, -4.5))
    D = float(10 ** rng.uniform(-9.7, -8.7))
    rho = float(10 ** rng.uniform(-6, -4)) if rng.random() < 0.5 else None
    kappa = float(10 ** rng.uniform(-5.5, -3.5)) if rng.random() < 0.4 else None
    if kind == "cyl":
        g = ds.Cylinder(R, (0.0, 0.0, 1.0), surface_relaxivity_t2=rho, permeability=kappa)
    elif kind == "sph":
        g = ds.Sphere(R, surface_relaxivity_t2=rho, permeability=kappa)
    elif kind == "ell":
        g = ds.Ellipsoid(R * rng.uniform(0.4, 1.0, 3), surface_relaxivity_t2=rho, permeability=kappa)
    else:
        r = R * rng.uniform(0.5, 1.0, 6)
        c, L, _ = ds.pack_dense(r, 0.4, seed=int(rng.integers(1 << 20)), dim=2)
        g = ds.PackedCylinders(r, c, L, surface_relaxivity_t2=rho, permeability=kappa)
    step = R * float(10 ** rng.uniform(np.log10(1 / 40), 0.0))
    dt = step ** 2 / (6 * D)
    return g, D, dt


@pytest.mark.parametrize("kind", ["cyl", "sph", "ell", "pcyl"])
def test_random_walks_bit_exact(gpu, oracle, kind):
    rng = np.random.default_rng({"cyl": 1, "sph": 2, "ell": 3, "pcyl": 4}[kind])
    for trial in range(12):
        g, D, dt = _random_case(rng, kind)
        n_t = int(rng.integers(5, 25))
        G = rng.normal(0, 0.05, (3, n_t, 3)).astype(np.float32)
        wf = ds.Waveform(G=G, dt=dt, echo_idx=n_t - 1)
        T2 = float(rng.uniform(0.02, 0.1)) if rng.random() < 0.3 else None
        seed = int(rng.integers(1 << 30))
        plan = ds.make_walk_plan(g, D, dt, T2=T2)
        # odd trials take the matrix-core phase update (32-measurement block; the random
        # G has a nonzero net moment, so the start point's phase is exercised too)
        mb = 32 if trial % 2 else 0
        print(f"{kind} trial {trial}: n_sub={plan.n_sub} wall={plan.wall} n_t={n_t} meas_block={mb}",
              flush=True)
        r = gpu.walk(g, D, wf, 256, seed=seed, T2=T2, outputs=("final_pos", "comp_final", "n_cross",
                                                                "log_w", "phi"), meas_block=mb)
        o = oracle.run_oracle(g, D, wf, 256, seed=seed, T2=T2)
        np.testing.assert_array_equal(r["final_pos"], o["final_pos"], err_msg=f"trial {trial}")
        np.testing.assert_array_equal(r["comp_final"], o["comp_final"])
        np.testing.assert_array_equal(r["n_cross"], o["n_cross"])
        np.testing.assert_array_equal(r["log_w"], o["log_w"])
        assert np.max(np.abs(r["signal"] - o["signal"])) <= 1e-5
        # per-walker phases: both sides sum n_t x n_sub float32 terms in different orders
        # (hundreds of sub-steps of absolute positions here), so the bar is relative to
        # the largest phase; a layout or start-point error would be of the phase's order
        bar = 2e-4 + 1e-3 * float(np.abs(o["phi"]).max())
        np.testing.assert_allclose(r["phi"], o["phi"], atol=bar, rtol=0, err_msg=f"trial {trial}")
