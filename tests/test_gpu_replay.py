"""Walk once, replay many: the GPU producer against the walk oracle (bit-exact
float16 positions and boundary log-weights) and the GPU replay contraction
against the float64 replay oracle (oracle/replay_oracle.py)."""
import numpy as np
import pytest

import dmipy_sim_amd as ds
from dmipy_sim_amd import trajectories as tj
from oracle import replay_oracle as ro
from substrates import D, packed_cyl

pytestmark = pytest.mark.gpu


def _producer_oracle(oracle, geom, n, seed, T_max, dt, save_relax):
    """The oracle walk with the producer's constants, per-sample outputs."""
    pp = tj.producer_plan(geom, D, T_max, dt, save_relax)
    p = oracle.oracle_plan(geom, D, pp["dt_actual"])
    p.update(n_sub=pp["sub_steps"], step_len=pp["step_len"], wall=pp["wall"],
             rho_over_D=pp["rho_over_D"], kappa_over_D=pp["kappa_over_D"],
             dt_sub=float(np.float32(pp["dt_sim"])), gamma_dt_sub=0.0, has_t2=False,
             has_t1=False, inv_T2=0.0, inv_T1=0.0)
    wf = ds.Waveform(G=np.zeros((1, pp["n_t"], 3), np.float32), dt=pp["dt_actual"],
                     echo_idx=pp["n_t"] - 1)
    o = oracle.run_oracle(geom, D, wf, n, seed=seed, plan=p, per_sample=True)
    return o, pp


GEOMS = [
    ("free", lambda: ds.FreeDiffusion()),
    ("box", lambda: ds.Box1D(6e-6)),
    ("cyl_rho", lambda: ds.Cylinder(3e-6, (0, 0, 1), surface_relaxivity_t2=2e-5)),
    ("cyl_rotated", lambda: ds.Cylinder(3e-6, (0.6, 0.0, 0.8))),
    ("sph", lambda: ds.Sphere(2.5e-6, surface_relaxivity_t2=1e-5)),
    ("sph_perm", lambda: ds.Sphere(3e-6, permeability=5e-5)),
    ("pcyl", lambda: packed_cyl(rho=1e-5)),
    ("ell_rho", lambda: ds.Ellipsoid((3e-6, 2e-6, 4e-6), surface_relaxivity_t2=1e-5)),
    ("ell_perm", lambda: ds.Ellipsoid((3e-6, 2e-6, 4e-6), permeability=3e-5)),
    ("slab", lambda: ds.PermeableSlab1D(8e-6, 3e-5)),
    ("shell", lambda: ds.PermeableShell(3e-6, 6e-6, 3e-5)),
    ("shell_cyl", lambda: ds.PermeableShell(3e-6, 6e-6, 3e-5, kind="cylinder",
                                            orientation=(0.6, 0.0, 0.8))),
]
PERMEABLE = ("sph_perm", "slab", "shell", "shell_cyl")


@pytest.mark.parametrize("name,factory", GEOMS, ids=[g[0] for g in GEOMS])
@pytest.mark.parametrize("save_relax", [False, True], ids=["positions", "relax"])
def test_producer_bit_exact(gpu, oracle, name, factory, save_relax):
    geom = factory()
    T_max, dt, n, seed = 0.012, 0.012 / 39, 300, 5
    print(f"producer {name} relax={save_relax}", flush=True)
    ts = gpu.TrajectorySet.simulate(n, D, geom, T_max, dt, seed=seed,
                                    save_relaxation_data=save_relax)
    traj, dlog, comp = ts.download()
    ts.close()
    o, pp = _producer_oracle(oracle, geom, n, seed, T_max, dt, save_relax)
    pos = o["pos_full"]
    plan = ds.make_walk_plan(geom, D, pp["dt_actual"])
    if plan.rot is not None:
        M = np.asarray(plan.rot, np.float64).T.astype(np.float32)
        x, y, z = pos[..., 0], pos[..., 1], pos[..., 2]
        pos = np.stack([(M[i, 0] * x + M[i, 1] * y) + M[i, 2] * z for i in range(3)], axis=-1)
    assert traj.dtype == np.float16 and traj.shape == (n, pp["n_t"], 3)
    np.testing.assert_array_equal(traj.view(np.uint16), pos.astype(np.float16).view(np.uint16))
    if save_relax:
        np.testing.assert_array_equal(dlog.view(np.uint16),
                                      o["dlog_full"].astype(np.float16).view(np.uint16))
        if name == "ell_perm":                  # permeable, no radius: float16 zeros
            assert comp.dtype == np.float16 and not comp.any()
        elif name in PERMEABLE:
            assert comp.dtype == np.float16
            np.testing.assert_array_equal(comp.view(np.uint16),
                                          o["occ_full"].astype(np.float16).view(np.uint16))
            assert 0 < comp.astype(float).mean() < 1
        else:
            assert comp.dtype == np.int8 and not comp.any()
        if name in ("cyl_rho", "sph", "pcyl", "ell_rho"):
            assert (dlog < 0).any()


@pytest.mark.parametrize("n_meas", [3, 20, 64, 150])
def test_replay_matches_oracle(gpu, n_meas):
    rng = np.random.default_rng(n_meas)
    N, n_t, dt = 1000 + n_meas, 90, 2e-4
    traj = rng.normal(0, 4e-6, (N, n_t, 3)).astype(np.float16)
    dlog = (-np.abs(rng.normal(0, 3e-8, (N, n_t)))).astype(np.float16)
    G = rng.normal(0, 0.04, (n_meas, n_t, 3)).astype(np.float32)
    chi = np.ones(n_t)
    chi[30:45] = 0
    print(f"replay n_meas={n_meas}", flush=True)
    for kw in (dict(), dict(chi_perp=chi, T2=0.07, T1=0.9),
               dict(chi_perp=chi, T2=0.07, surface_relaxivity=3e-5, D=D)):
        dl = dlog if "surface_relaxivity" in kw else None
        phi, lw, S = gpu.replay(traj, dt, G, dt, dlog_boundary_unit=dl,
                                return_walker_signals=True, **kw)
        So, phio, lwo = ro.replay(traj, dt, G, dt, chi=kw.get("chi_perp"), dlog=dl,
                                  T2=kw.get("T2"), T1=kw.get("T1"),
                                  rho=kw.get("surface_relaxivity"), D=kw.get("D"))
        np.testing.assert_allclose(S, So, rtol=0, atol=1e-10)
        np.testing.assert_allclose(phi, phio, rtol=1e-10, atol=1e-10 * np.max(np.abs(phio)))
        np.testing.assert_allclose(lw, np.broadcast_to(lwo, lw.shape), rtol=1e-12, atol=1e-14)


def test_replay_per_measurement_chi_resampling_and_stimulated_echo(gpu):
    rng = np.random.default_rng(7)
    N, n_t, dt = 700, 80, 1e-4
    traj = rng.normal(0, 3e-6, (N, n_t, 3)).astype(np.float16)
    dlog = (-np.abs(rng.normal(0, 3e-8, (N, n_t)))).astype(np.float16)
    n_wf, dt_wf = 41, 2e-4                        # coarser waveform grid: linear resampling
    G = rng.normal(0, 0.05, (5, n_wf, 3)).astype(np.float32)
    chi2 = (rng.random((5, n_wf)) > 0.3).astype(float)
    S = gpu.replay(traj, dt, G, dt_wf, chi_perp=chi2, dlog_boundary_unit=dlog,
                   surface_relaxivity=2e-5, D=D, T2=0.05, stimulated_echo=True)
    So, _, _ = ro.replay(traj, dt, G, dt_wf, chi=chi2, dlog=dlog, T2=0.05, rho=2e-5, D=D,
                         stimulated_echo=True)
    np.testing.assert_allclose(S, So, rtol=0, atol=1e-10)


def test_replay_of_producer_trajectories_resident(gpu):
    """Replaying a device-resident trajectory set equals replaying its download."""
    geom = ds.Cylinder(4e-6, (0, 0, 1), surface_relaxivity_t2=1e-5)
    wf = ds.set_b(ds.pgse(0.004, 0.01, 0.05, [[1, 0, 0], [0, 0, 1]], 60), [1e9, 1e9])
    ts = gpu.TrajectorySet.simulate(2000, D, geom, wf.dt * 59, wf.dt, seed=3,
                                    save_relaxation_data=True)
    S_dev = ts.replay(wf.G, wf.dt, surface_relaxivity=1e-5, D=D)
    traj, dlog, _ = ts.download()
    ts.close()
    So, _, _ = ro.replay(traj, wf.dt, wf.G, wf.dt, dlog=dlog, rho=1e-5, D=D)
    np.testing.assert_allclose(S_dev, So, rtol=0, atol=1e-10)
    assert 0.1 < S_dev[1] < S_dev[0] < 1.0      # free along the axis, restricted across it


def test_simulate_replay_engine_physics(gpu):
    """engine='replay': free diffusion at exp(-bD) within 3 SE, and agreement with the
    fused walk on a relaxing cylinder within 4 combined SE."""
    wf = ds.set_b(ds.pgse(0.01, 0.04, 0.05, [[1, 0, 0]] * 3, 300), [0.0, 1e9, 2e9])
    S = gpu.simulate(50_000, D, wf, ds.FreeDiffusion(), seed=0, engine="replay")
    ref = np.exp(-np.array([0.0, 1e9, 2e9]) * D)
    se = np.sqrt(np.maximum((1 + np.exp(-4 * np.array([0.0, 1e9, 2e9]) * D)) / 2 - ref ** 2, 1e-12)
                 / 50_000)
    assert S.dtype == np.float32
    assert np.all(np.abs(S - ref) <= 3 * se + 1e-6)
    geom = ds.Cylinder(5e-6, (0, 0, 1), surface_relaxivity_t2=1e-6)
    wf2 = ds.set_b(ds.pgse(0.01, 0.04, 0.05, [[1, 0, 0], [0.6, 0, 0.8]], 400), [2e9, 2e9])
    Sr = gpu.simulate(200_000, D, wf2, geom, seed=1, engine="replay")
    Sf = gpu.simulate(200_000, D, wf2, geom, seed=2, engine="fused")
    se2 = np.sqrt(2 * 0.5 / 200_000)
    assert np.all(np.abs(Sr - Sf) <= 4 * se2)
