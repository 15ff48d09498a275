"""Batched persistent launch vs the CPU oracle, instance by instance.

Every instance of a batch walks the Philox stream with its own instance word, so
instance i must reproduce the oracle run with ``instance=i`` and the same seed:
the same trajectories, signals within 1e-5 (phase summation order differs, as in
the single-walk parity tests).  The batch mixes every geometry kind and wall
mode, ragged walker counts, two waveform sets and a rotated cylinder (which gets
its own rotated set).
"""
import numpy as np
import pytest

import dmipy_sim_amd as ds
from substrates import CASES, D, small_pgse

pytestmark = pytest.mark.gpu


def _second_waveform():
    wf = small_pgse(b=(0.5e9, 1.5e9, 2.5e9, 4e9))
    G = wf.G[:, :, [2, 0, 1]].copy()   # different directions, same (n_meas, n_t)
    return ds.Waveform(G=G, dt=wf.dt, echo_idx=wf.echo_idx)


def _instances():
    wfs = [small_pgse(), _second_waveform()]
    specs = []
    counts = [64, 100, 257, 128, 65, 192, 130, 63]
    for k, (name, factory, T2, _) in enumerate(CASES):
        specs.append(dict(name=name, geometry=factory(), diffusivity=D * (1.0 + 0.1 * (k % 3)),
                          n_walkers=counts[k % len(counts)], T2=T2, waveform=k % 2))
    specs.append(dict(name="cyl_rotated", geometry=ds.Cylinder(3e-6, (1.0, 0.0, 0.0)),
                      diffusivity=D, n_walkers=150, T2=None, waveform=0))
    return wfs, specs


def test_batch_matches_oracle_per_instance(gpu, oracle):
    wfs, specs = _instances()
    seed = 77
    print(f"batch: {len(specs)} instances, {len(wfs)} waveforms", flush=True)
    bw = gpu.BatchedWalk(specs, wfs)
    assert bw.n_sets == 3      # two waveforms + the rotated copy of the first
    sig, stats = bw.run(seed)
    bw.close()
    assert stats["walker_steps"] > 0 and stats["n_units"] == sum(
        (s["n_walkers"] + 63) // 64 for s in specs)
    for i, s in enumerate(specs):
        wf = wfs[s["waveform"]]
        plan = ds.make_walk_plan(s["geometry"], s["diffusivity"], wf.dt, T2=s["T2"])
        G = wf.G
        if plan.rot is not None:
            G = np.einsum('ij,mtj->mti', plan.rot.astype(np.float64), G).astype(np.float32)
        o = oracle.run_oracle(s["geometry"], s["diffusivity"], wf, s["n_walkers"], seed=seed,
                              T2=s["T2"], G_native=G, instance=i)
        err = np.max(np.abs(sig[i] - o["signal"]))
        assert err <= 1e-5, f"instance {i} ({s['name']}): max |dS| = {err:.3g}"


@pytest.mark.parametrize("n_dirs", [10, 23])
def test_batch_screened_packings_parked_phase_match_oracle(gpu, oracle, n_dirs):
    """Permeable packings walk in the screened kernel, and with more than 16
    measurements its phase is parked in device memory between samples: 40
    measurements (the 64-wide block, mixed with an instance of the general
    kernel: both kernels in one graph) and 92 (the 128-wide block, which only
    the screened kernel takes).  Every instance matches the oracle, and a graph
    replay repeats the signals exactly."""
    from substrates import packed_cyl, packed_sph, tight_cell_cyl
    wf = small_pgse(n_dirs=n_dirs)
    n_meas = wf.G.shape[0]
    assert n_meas == 4 * n_dirs
    specs = [dict(geometry=packed_cyl(kappa=5e-5, rho=1e-5), diffusivity=D, n_walkers=300),
             dict(geometry=packed_sph(kappa=5e-5), diffusivity=D, n_walkers=200),
             dict(geometry=packed_cyl(kappa=2e-5, seed=7), diffusivity=1.5 * D, n_walkers=257),
             dict(geometry=tight_cell_cyl(), diffusivity=D, n_walkers=150),   # whole-list path
             # 24 spheres at VF 0.5: 46 cells with shrunk boxes, one at level 2
             dict(geometry=packed_sph(kappa=5e-5, rho=1e-5, n=24, vf=0.5, seed=6), diffusivity=D,
                  n_walkers=400)]
    if n_meas <= 64:
        specs.insert(2, dict(geometry=ds.Cylinder(3e-6, (0, 0, 1), permeability=5e-5), diffusivity=D,
                             n_walkers=130))
    seed = 31
    print(f"batch, screened kernel: {len(specs)} instances, {n_meas} measurements, seed {seed}", flush=True)
    bw = gpu.BatchedWalk(specs, wf)
    sig, _ = bw.run(seed)
    sig2, _ = bw.run(seed)
    bw.close()
    np.testing.assert_array_equal(sig, sig2)
    for i, s in enumerate(specs):
        o = oracle.run_oracle(s["geometry"], s["diffusivity"], wf, s["n_walkers"], seed=seed, instance=i)
        err = np.max(np.abs(sig[i] - o["signal"]))
        assert err <= 1e-5, f"instance {i}: max |dS| = {err:.3g}"


def test_batch_instance0_matches_single_walk(gpu):
    wf = small_pgse()
    geom = ds.Cylinder(3e-6, (0, 0, 1), permeability=5e-5)
    specs = [dict(geometry=geom, diffusivity=D, n_walkers=300),
             dict(geometry=ds.Sphere(2e-6), diffusivity=D, n_walkers=100)]
    bw = gpu.BatchedWalk(specs, wf)
    sig, _ = bw.run(9)
    bw.close()
    one = gpu.walk(geom, D, wf, 300, seed=9)
    assert np.max(np.abs(sig[0] - one["signal"])) <= 1e-5


def test_batch_graph_replays_and_seeds(gpu):
    """Graph replays give identical results, equal to the un-captured launch;
    a new seed changes the result."""
    wf = small_pgse()
    specs = [dict(geometry=ds.Sphere(3e-6 + 1e-7 * k), diffusivity=D, n_walkers=256)
             for k in range(40)]
    specs += [dict(geometry=CASES[10][1](), diffusivity=D, n_walkers=200)]
    g = gpu.BatchedWalk(specs, wf, use_graph=True)
    e = gpu.BatchedWalk(specs, wf, use_graph=False)
    a1, _ = g.run(3)
    ref, _ = e.run(3)
    a2, _ = g.run(3)
    a3, _ = g.run(3)
    b1, _ = g.run(4)
    g.close()
    e.close()
    np.testing.assert_array_equal(a1, ref)
    np.testing.assert_array_equal(a1, a2)
    np.testing.assert_array_equal(a1, a3)
    assert not np.array_equal(a1, b1)


def test_batch_chi_schedule(gpu, oracle):
    """chi_perp of the waveforms gates wall relaxation and T2 in every instance."""
    wf = small_pgse()
    chi = np.ones(wf.G.shape[1], np.float32)
    chi[10:18] = 0.0
    wf.chi_perp = chi
    specs = [dict(geometry=ds.Sphere(3e-6, surface_relaxivity_t2=2e-5), diffusivity=D,
                  n_walkers=192, T2=0.05),
             dict(geometry=CASES[11][1](), diffusivity=D, n_walkers=128, T2=0.05)]
    bw = gpu.BatchedWalk(specs, wf)
    sig, _ = bw.run(21)
    bw.close()
    for i, s in enumerate(specs):
        o = oracle.run_oracle(s["geometry"], D, wf, s["n_walkers"], seed=21, T2=0.05, instance=i)
        assert np.max(np.abs(sig[i] - o["signal"])) <= 1e-5


def test_simulate_mixture_equals_serial_calls(gpu):
    """simulate_mixture runs every compartment in one batched launch; compartment
    i walks the trajectories of simulate(..., seed=seed+i) (core.py:821-858)."""
    wf = small_pgse()
    comps = [dict(fraction=0.4, n_walkers=500, diffusivity=D, geometry=ds.Cylinder(3e-6, (0, 0, 1))),
             dict(fraction=0.3, n_walkers=300, diffusivity=1.5e-9, geometry=ds.Sphere(4e-6)),
             dict(fraction=0.2, n_walkers=257, diffusivity=3e-9, geometry=ds.FreeDiffusion()),
             dict(fraction=0.1, n_walkers=200, diffusivity=D, geometry=CASES[12][1]())]
    mix = gpu.simulate_mixture(comps, wf, seed=31)
    ref = sum(c["fraction"] * gpu.simulate(c["n_walkers"], c["diffusivity"], wf, c["geometry"],
                                           seed=31 + i)
              for i, c in enumerate(comps))
    assert mix.dtype == np.float32 and mix.shape == (wf.G.shape[0],)
    assert np.max(np.abs(mix - ref)) <= 1e-5


def test_simulate_mixture_with_myelinated_compartment(gpu):
    """A MyelinatedCylinder compartment (diffusivity None, its own per-compartment
    diffusivities) cannot ride the batched launch: the mixture walks every
    compartment with simulate(..., seed=seed+i) instead (core.py:846-856) and
    equals those serial calls exactly."""
    wf = small_pgse()
    myel = ds.MyelinatedCylinder(3e-6, 5e-6, (0, 0, 1), D_intra=2e-9, D_extra=1.5e-9,
                                 D_myelin=4e-10, T2_myelin=0.015)
    comps = [dict(fraction=0.6, n_walkers=900, diffusivity=None, geometry=myel),
             dict(fraction=0.4, n_walkers=500, diffusivity=D, geometry=ds.Cylinder(3e-6, (0, 0, 1)))]
    print("mixture: MyelinatedCylinder + Cylinder, 16 measurements", flush=True)
    mix = gpu.simulate_mixture(comps, wf, seed=40)
    ref = None
    for i, c in enumerate(comps):
        s = c["fraction"] * gpu.simulate(c["n_walkers"], c["diffusivity"], wf, c["geometry"], seed=40 + i)
        ref = s if ref is None else ref + s
    assert mix.dtype == np.float32 and mix.shape == (wf.G.shape[0],)
    np.testing.assert_array_equal(mix, ref)
    assert np.all(mix[wf.G.reshape(wf.G.shape[0], -1).any(1)] < 1.0)


def test_launch_shapes_give_identical_signals(gpu, oracle):
    """Every launch shape (persistent waves per SIMD, walkers per wavefront pop,
    sub-step unrolling) walks the same tiles with the same Philox streams and
    writes the same partial rows: the signals are bit-identical, and the
    unrolled kernel still reproduces the oracle."""
    wfs, specs = _instances()
    ref = None
    for shape in ((0, 0, 0), (4, 512, 2), (2, 128, 1), (1, 256, 2), (3, 64, 1)):
        print(f"batch launch shape {shape}", flush=True)
        bw = gpu.BatchedWalk(specs, wfs, waves_per_simd=shape[0], walkers_per_wave=shape[1],
                             unroll=shape[2])
        sig, _ = bw.run(31)
        bw.close()
        if ref is None:
            ref = sig
        assert np.array_equal(sig, ref), shape
    s = specs[3]
    o = oracle.run_oracle(s["geometry"], s["diffusivity"], wfs[s["waveform"]], s["n_walkers"],
                          seed=31, T2=s["T2"], instance=3)
    assert np.max(np.abs(ref[3] - o["signal"])) <= 1e-5


def test_autotuner_times_every_knob_and_caches(gpu, tmp_path, monkeypatch):
    """tune_batch times the three knobs by coordinate descent on the batch and
    keeps the fastest under the batch's key; tuning='auto' then takes it."""
    from dmipy_sim_amd import autotune
    monkeypatch.setenv("DMIPY_AUTOTUNE_CACHE", str(tmp_path / "tune.json"))
    autotune.clear()
    autotune._loaded["done"] = False
    wfs, specs = _instances()
    best, times = autotune.tune_batch(specs, wfs, walker_scale=1.0, reps=1,
                                      log=lambda s: print(s, flush=True))
    assert times[best] == min(times.values())
    seen = {k[0] for k in times}, {k[1] for k in times}, {k[2] for k in times}
    assert seen[0] == set(autotune.BATCH_WAVES_PER_SIMD)
    assert seen[1] == set(autotune.BATCH_WALKERS_PER_WAVE)
    assert seen[2] == set(autotune.BATCH_UNROLL)
    assert (tmp_path / "tune.json").exists()
    bw = gpu.BatchedWalk(specs, wfs, tuning="auto")
    assert (bw.shape["waves_per_simd"], bw.shape["walkers_per_wave"], bw.shape["unroll"]) == best
    bw.close()
    autotune.clear()
