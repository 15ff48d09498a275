"""GPU walk vs CPU oracle on identical inputs (bit-exact positions and counts).

With the same injected normal triples and Powles uniforms, the HIP kernel and the
oracle's restatement of the reference step (walk_oracle.c) must produce identical
float32 positions, log-weights, compartments and crossing counts; signals differ
only by the phase summation order (per sample on the GPU, per sub-step in the
oracle) and must agree within 1e-5.  With the Philox stream (no injection) both
sides draw the same numbers, so the same bar applies.
"""
import numpy as np
import pytest

import dmipy_sim_amd as ds
from substrates import WALK_CASES, D, small_pgse, inside_starts

pytestmark = pytest.mark.gpu

OUTS = ("final_pos", "comp_origin", "comp_final", "n_cross", "log_w", "phi")


def _run_pair(gpu, oracle, geom, T2, T1, n, seed, inject, chi=None, kahan=False, meas_block=0):
    wf = small_pgse()
    if chi is not None:
        wf.chi_perp = chi
    plan = ds.make_walk_plan(geom, D, wf.dt, T2=T2, T1=T1)
    S = wf.G.shape[1] * plan.n_sub
    r0 = inside_starts(geom, n, seed)
    kw = {}
    if inject:
        rng = np.random.default_rng(seed)
        kw["inject_normals"] = rng.standard_normal((n, S, 3)).astype(np.float32)
        kw["inject_uniforms"] = rng.random((n, S)).astype(np.float32)
    print(f"parity: {type(geom).__name__} wall={plan.wall} n_sub={plan.n_sub} N={n} "
          f"inject={inject}", flush=True)
    g = gpu.walk(geom, D, wf, n, seed=seed, T2=T2, T1=T1, r0=r0, outputs=OUTS + ("signal_im",),
                 kahan=kahan, meas_block=meas_block, **kw)
    o = oracle.run_oracle(geom, D, wf, n, seed=seed, T2=T2, T1=T1, r0=r0, **kw)
    return g, o, plan


@pytest.mark.parametrize("name,factory,T2,T1", WALK_CASES, ids=[c[0] for c in WALK_CASES])
@pytest.mark.parametrize("inject", [True, False], ids=["injected", "philox"])
def test_walk_bit_exact(gpu, oracle, name, factory, T2, T1, inject):
    geom = factory()
    g, o, plan = _run_pair(gpu, oracle, geom, T2, T1, n=384, seed=11, inject=inject)
    np.testing.assert_array_equal(g["final_pos"], o["final_pos"])
    np.testing.assert_array_equal(g["comp_origin"], o["comp_origin"])
    np.testing.assert_array_equal(g["comp_final"], o["comp_final"])
    np.testing.assert_array_equal(g["n_cross"], o["n_cross"])
    np.testing.assert_array_equal(g["log_w"], o["log_w"])
    assert np.max(np.abs(g["signal"] - o["signal"])) <= 1e-5
    assert np.max(np.abs(g["signal_im"] - o["signal_im"])) <= 1e-5
    np.testing.assert_allclose(g["phi"], o["phi"], atol=2e-4, rtol=1e-5)
    if plan.wall == 2 or name == "myel_perm":
        assert g["n_cross"].sum() > 0, "permeable case produced no crossings"


def test_tight_cell_whole_list_path(gpu, oracle):
    """A permeable packing whose cell is too small for cached nearest images: the
    screened walk tests the whole cell list every sub-step, minimum images by true
    division; bit-exact against the oracle."""
    from substrates import tight_cell_cyl
    geom = tight_cell_cyl()
    g, o, plan = _run_pair(gpu, oracle, geom, None, None, n=320, seed=5, inject=False)
    assert plan.wall == 2
    for k in ("final_pos", "comp_origin", "comp_final", "n_cross", "log_w"):
        np.testing.assert_array_equal(g[k], o[k], err_msg=k)
    assert np.max(np.abs(g["signal"] - o["signal"])) <= 1e-5
    assert g["n_cross"].sum() > 0


def test_stored_magnetisation_chi_schedule(gpu, oracle):
    """chi_perp gates wall relaxation and T2 (transverse) against T1 (stored)."""
    wf = small_pgse()
    chi = np.ones(wf.G.shape[1], np.float32)
    chi[10:20] = 0.0
    geom = ds.Sphere(3e-6, surface_relaxivity_t2=2e-5)
    g, o, _ = _run_pair(gpu, oracle, geom, 0.05, 0.8, 256, 5, True, chi=chi)
    np.testing.assert_array_equal(g["log_w"], o["log_w"])
    assert np.max(np.abs(g["signal"] - o["signal"])) <= 1e-5


@pytest.mark.parametrize("mb", [4, 16, 32, 64, 128])
def test_measurement_blocks_agree(gpu, oracle, mb):
    """Every measurement-block width gives the oracle's signal (walk recomputed per block)."""
    geom = ds.Cylinder(3e-6, (0, 0, 1), surface_relaxivity_t2=2e-5)
    g, o, _ = _run_pair(gpu, oracle, geom, None, None, 256, 3, True, meas_block=mb)
    assert np.max(np.abs(g["signal"] - o["signal"])) <= 1e-5
    np.testing.assert_array_equal(g["final_pos"], o["final_pos"])


@pytest.mark.parametrize("kahan", [False, True])
def test_kahan_phase(gpu, oracle, kahan):
    geom = packed_geom = None
    from substrates import packed_cyl
    geom = packed_cyl(kappa=5e-5)
    g, o, _ = _run_pair(gpu, oracle, geom, None, None, 256, 9, True, kahan=kahan)
    assert np.max(np.abs(g["signal"] - o["signal"])) <= 1e-5


def test_rotated_cylinder_matches_native_frame(gpu):
    """A cylinder along x with the gradient rotated accordingly gives the same signal
    as the z cylinder (native-frame walk; identical Philox stream)."""
    wf = small_pgse()
    gz = ds.Cylinder(3e-6, (0, 0, 1), surface_relaxivity_t2=2e-5)
    gx = ds.Cylinder(3e-6, (1, 0, 0), surface_relaxivity_t2=2e-5)
    Rz_to_x = gx._R_inv.astype(np.float64)   # native z axis -> lab x axis
    wf_x = ds.Waveform(G=np.einsum('ij,mtj->mti', Rz_to_x, wf.G).astype(np.float32), dt=wf.dt,
                       echo_idx=wf.echo_idx)
    a = gpu.walk(gz, D, wf, 4096, seed=2, outputs=("final_pos",))
    b = gpu.walk(gx, D, wf_x, 4096, seed=2, outputs=("final_pos",))
    assert np.max(np.abs(a["signal"] - b["signal"])) <= 2e-5
    np.testing.assert_allclose(b["final_pos"], (Rz_to_x @ a["final_pos"].T.astype(np.float64)).T,
                               atol=1e-11)


def test_determinism(gpu):
    geom = ds.Sphere(3e-6, permeability=5e-5)
    wf = small_pgse()
    a = gpu.walk(geom, D, wf, 5000, seed=42, outputs=("final_pos", "n_cross"))
    b = gpu.walk(geom, D, wf, 5000, seed=42, outputs=("final_pos", "n_cross"))
    np.testing.assert_array_equal(a["signal"], b["signal"])
    np.testing.assert_array_equal(a["final_pos"], b["final_pos"])
    c = gpu.walk(geom, D, wf, 5000, seed=43)
    assert not np.array_equal(a["signal"], c["signal"])


def test_ragged_walker_counts(gpu, oracle):
    """Walker counts that do not fill a wavefront / workgroup (1, 63, 65, 257)."""
    geom = ds.Cylinder(3e-6, (0, 0, 1), permeability=5e-5)
    for n in (1, 63, 65, 257):
        g, o, _ = _run_pair(gpu, oracle, geom, None, None, n, 21, True)
        np.testing.assert_array_equal(g["final_pos"], o["final_pos"])
        assert np.max(np.abs(g["signal"] - o["signal"])) <= 1e-5


def test_cpmg_echoes_match_oracle(gpu, oracle):
    wf = ds.cpmg(6, 10e-3, 0.02, [[1, 0, 0], [0, 1, 0]], n_t_per_echo=20)
    geom = ds.Sphere(4e-6, surface_relaxivity_t2=1e-5)
    n = 512
    r0 = geom.init_positions(n, 1)
    plan = ds.make_walk_plan(geom, D, wf.dt, T2=0.07)
    rng = np.random.default_rng(1)
    nrm = rng.standard_normal((n, wf.G.shape[1] * plan.n_sub, 3)).astype(np.float32)
    g = gpu.walk(geom, D, wf, n, T2=0.07, r0=r0, inject_normals=nrm, echo_indices=wf.echo_indices)
    o = oracle.run_oracle(geom, D, wf, n, T2=0.07, r0=r0, inject_normals=nrm,
                          echo_indices=wf.echo_indices)
    assert g["signal"].shape == (6, 2)
    assert np.max(np.abs(g["signal"] - o["signal"])) <= 1e-5


def _unbalanced(wf, factor=1.5):
    """The waveform with its first half scaled: a nonzero net moment sum_t G."""
    G = wf.G.copy()
    G[:, : G.shape[1] // 2] *= factor
    return ds.Waveform(G=G, dt=wf.dt, echo_idx=wf.echo_idx)


@pytest.mark.parametrize("mb", [32, 128])
@pytest.mark.parametrize("where", ["cylinder_axis", "packed_far_cell"])
def test_matrix_phase_far_start_net_moment(gpu, oracle, mb, where):
    """Matrix-core phase (float16 hi/lo split of displacement sums) with start points
    far from the origin and a waveform whose net moment is not zero, so the start
    point's own phase (start . sum_t G) is a large part of every walker's phase."""
    wf = _unbalanced(small_pgse())
    if where == "cylinder_axis":
        geom = ds.Cylinder(3e-6, (0, 0, 1), surface_relaxivity_t2=2e-5)
        r0 = inside_starts(geom, 256, 7)
        r0[:, 2] += 2e-5                      # 20 um along the free axis
    else:
        from substrates import packed_cyl
        geom = packed_cyl(kappa=5e-5)
        r0 = inside_starts(geom, 256, 7)
        r0[:, 0] += 3 * geom._L_float         # three cells away: cell index 3 after the fold
    plan = ds.make_walk_plan(geom, D, wf.dt)
    rng = np.random.default_rng(7)
    S = wf.G.shape[1] * plan.n_sub
    kw = {"inject_normals": rng.standard_normal((256, S, 3)).astype(np.float32),
          "inject_uniforms": rng.random((256, S)).astype(np.float32)}
    print(f"matrix phase: {where} mb={mb}", flush=True)
    g = gpu.walk(geom, D, wf, 256, r0=r0, outputs=("final_pos", "phi", "signal_im"), meas_block=mb, **kw)
    o = oracle.run_oracle(geom, D, wf, 256, r0=r0, **kw)
    np.testing.assert_array_equal(g["final_pos"], o["final_pos"])
    assert np.abs(o["phi"]).max() > 3.0, "start-point phase too small to test"
    np.testing.assert_allclose(g["phi"], o["phi"], atol=2e-4, rtol=1e-5)
    assert np.max(np.abs(g["signal"] - o["signal"])) <= 1e-5
    assert np.max(np.abs(g["signal_im"] - o["signal_im"])) <= 1e-5


def test_cpmg_echoes_on_matrix_cores(gpu, oracle):
    """Echo rows emitted from the matrix-core accumulators (32-measurement block)."""
    wf = ds.cpmg(6, 10e-3, 0.02, [[1, 0, 0], [0, 1, 0]], n_t_per_echo=20)
    geom = ds.Sphere(4e-6, surface_relaxivity_t2=1e-5)
    n = 512
    r0 = geom.init_positions(n, 1)
    plan = ds.make_walk_plan(geom, D, wf.dt, T2=0.07)
    rng = np.random.default_rng(1)
    nrm = rng.standard_normal((n, wf.G.shape[1] * plan.n_sub, 3)).astype(np.float32)
    g = gpu.walk(geom, D, wf, n, T2=0.07, r0=r0, inject_normals=nrm, echo_indices=wf.echo_indices,
                 meas_block=32)
    o = oracle.run_oracle(geom, D, wf, n, T2=0.07, r0=r0, inject_normals=nrm,
                          echo_indices=wf.echo_indices)
    assert g["signal"].shape == (6, 2)
    assert np.max(np.abs(g["signal"] - o["signal"])) <= 1e-5


def test_large_tables_take_a_narrower_block(gpu, oracle):
    """2025 cylinders: the object table and neighbour grid leave too little LDS for the
    128-measurement block, so the plan takes a narrower one; results are unchanged."""
    k = 45
    h = 4.5e-6
    xs = (np.arange(k) + 0.5) * h
    centers = np.stack(np.meshgrid(xs, xs, indexing="ij"), -1).reshape(-1, 2)
    radii = np.full(len(centers), 1.6e-6)
    geom = ds.PackedCylinders(radii, centers, k * h, surface_relaxivity_t2=2e-5)
    wf = small_pgse(n_t=12, n_dirs=32)       # 128 measurements
    n = 256
    r0 = geom.init_positions(n, 5)
    plan = ds.make_walk_plan(geom, D, wf.dt)
    rng = np.random.default_rng(5)
    nrm = rng.standard_normal((n, wf.G.shape[1] * plan.n_sub, 3)).astype(np.float32)
    print(f"large tables: n_obj={len(radii)} n_sub={plan.n_sub}", flush=True)
    g = gpu.walk(geom, D, wf, n, r0=r0, inject_normals=nrm, outputs=("final_pos",))
    o = oracle.run_oracle(geom, D, wf, n, r0=r0, inject_normals=nrm)
    np.testing.assert_array_equal(g["final_pos"], o["final_pos"])
    assert np.max(np.abs(g["signal"] - o["signal"])) <= 1e-5


def test_many_measurements_span_several_passes(gpu, oracle):
    """300 measurements exceed the widest measurement block (128): the walk runs
    several passes over the same trajectories (the last one padded), and every
    measurement still matches the oracle within 1e-5 with positions bit-exact."""
    wf = small_pgse(n_t=24, n_dirs=75, b=(0.5e9, 1e9, 2e9, 3e9))
    assert wf.G.shape[0] == 300
    geom = ds.Cylinder(4e-6, (0, 0, 1), surface_relaxivity_t2=2e-6)
    n = 320
    r0 = geom.init_positions(n, 5)
    g = gpu.walk(geom, D, wf, n, seed=5, r0=r0, outputs=("final_pos", "log_w"))
    o = oracle.run_oracle(geom, D, wf, n, seed=5, r0=r0)
    np.testing.assert_array_equal(g["final_pos"], o["final_pos"])
    np.testing.assert_array_equal(g["log_w"], o["log_w"])
    assert g["signal"].shape == (300,)
    assert np.max(np.abs(g["signal"] - o["signal"])) <= 1e-5


def test_zero_gradient_and_two_samples(gpu):
    """No gradient and no relaxation: every weight is 1 and every phase 0, so the
    signal is exactly 1; a two-sample waveform walks (n_t = 2 is the least)."""
    G = np.zeros((3, 2, 3), np.float32)
    wf = ds.Waveform(G=G, dt=1e-3, echo_idx=1)
    for geom in (ds.Sphere(3e-6), ds.Cylinder(3e-6, (0, 0, 1)), ds.FreeDiffusion()):
        S = gpu.simulate(1000, D, wf, geom, seed=1)
        assert S.dtype == np.float32
        np.testing.assert_array_equal(S, np.ones(3, np.float32))
