"""Per-sample export (pos_full, comp_full) against the oracle's per-sample record.

The reference returns the position and compartment of every walker after every
sample when asked (core.py:579-617, 638-696).  The HIP kernel writes them once
per sample from the walker's folded offset and cell (positions unfolded, lab
frame); the oracle records its own walk at the same points.  On identical inputs
(injected increments and start points, or the same Philox stream) both must agree
bit for bit, at every sample, for packed substrates (periodic cells) and for a
rotated geometry (the lab-frame round trip), and the last sample must be the
final position / compartment.
"""
import numpy as np
import pytest

import dmipy_sim_amd as ds
from substrates import D, inside_starts, packed_cyl, packed_sph, small_pgse

pytestmark = pytest.mark.gpu

OUTS = ("final_pos", "comp_final", "pos_full", "comp_full", "n_cross")


@pytest.mark.parametrize("name,factory,inject", [
    ("pcyl_perm", lambda: packed_cyl(kappa=5e-5, rho=1e-5), True),
    ("pcyl_perm", lambda: packed_cyl(kappa=5e-5, rho=1e-5), False),
    ("psph_perm", lambda: packed_sph(kappa=5e-5), True),
    ("pcyl_reflect", lambda: packed_cyl(rho=2e-5), False),
], ids=["pcyl_perm-injected", "pcyl_perm-philox", "psph_perm-injected", "pcyl_reflect-philox"])
def test_packed_per_sample_export_bit_exact(gpu, oracle, name, factory, inject):
    geom = factory()
    wf = small_pgse()
    n = 384
    plan = ds.make_walk_plan(geom, D, wf.dt)
    S = wf.G.shape[1] * plan.n_sub
    r0 = inside_starts(geom, n, 7)
    kw = {}
    if inject:
        rng = np.random.default_rng(7)
        kw["inject_normals"] = rng.standard_normal((n, S, 3)).astype(np.float32)
        kw["inject_uniforms"] = rng.random((n, S)).astype(np.float32)
    print(f"export {name}: n_sub={plan.n_sub} N={n} n_t={wf.G.shape[1]} inject={inject}", flush=True)
    g = gpu.walk(geom, D, wf, n, seed=7, r0=r0, outputs=OUTS, **kw)
    o = oracle.run_oracle(geom, D, wf, n, seed=7, r0=r0, per_sample=True, **kw)
    assert g["pos_full"].shape == (n, wf.G.shape[1], 3)
    np.testing.assert_array_equal(g["pos_full"], o["pos_full"])
    np.testing.assert_array_equal(g["comp_full"], o["comp_full"])
    np.testing.assert_array_equal(g["pos_full"][:, -1], g["final_pos"])
    np.testing.assert_array_equal(g["comp_full"][:, -1], g["comp_final"])
    # compartments do change along the way on permeable walls
    if "perm" in name:
        assert np.any(g["comp_full"][:, 1:] != g["comp_full"][:, :-1])
        assert g["n_cross"].sum() > 0


@pytest.mark.parametrize("factory", [
    lambda: ds.Cylinder(3e-6, (0.6, 0.0, 0.8), permeability=5e-5, surface_relaxivity_t2=1e-5),
    lambda: ds.PackedCylinders(*_pack(), orientation=(0.0, 0.6, 0.8), permeability=5e-5),
], ids=["cylinder_perm", "pcyl_perm"])
def test_rotated_per_sample_export(gpu, oracle, factory):
    """Rotated geometry, Philox starts and steps (both sides walk in the native
    frame): the lab-frame export rotated back equals the oracle's native record."""
    geom = factory()
    wf = small_pgse()
    n = 512
    plan = ds.make_walk_plan(geom, D, wf.dt)
    assert plan.rot is not None
    G_native = np.einsum('ij,mtj->mti', plan.rot.astype(np.float64), wf.G).astype(np.float32)
    print(f"export rotated {type(geom).__name__}: n_sub={plan.n_sub} N={n}", flush=True)
    g = gpu.walk(geom, D, wf, n, seed=13, outputs=OUTS)
    o = oracle.run_oracle(geom, D, wf, n, seed=13, per_sample=True, G_native=G_native)
    back = np.einsum('ij,wtj->wti', plan.rot.astype(np.float64), g["pos_full"].astype(np.float64))
    np.testing.assert_allclose(back, o["pos_full"], rtol=0, atol=5e-12)
    np.testing.assert_array_equal(g["comp_full"], o["comp_full"])
    np.testing.assert_array_equal(g["pos_full"][:, -1], g["final_pos"])
    assert np.any(g["comp_full"][:, 1:] != g["comp_full"][:, :-1])


def _pack():
    rng = np.random.default_rng(5)
    radii = rng.uniform(2e-6, 3.5e-6, 8)
    c, L, _ = ds.pack_dense(radii, 0.45, seed=5, dim=2)
    return radii, c, L
