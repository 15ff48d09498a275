"""dmw_plan_advance: a walk split over several launches, the walkers kept in
device memory between them, equals the one-call walk bit for bit.

Every per-walker output (final and per-sample positions and compartments,
crossings, log-weights, phases), the signal, its imaginary part, the echo rows of
a CPMG walk and the wall-event counters are compared with dmw_walk on the same
seed, for slice boundaries that fall inside the waveform's lobes and on its echo
samples.  The cases cover the state the launches hand over: the periodic cell
(packed cylinders and spheres), the screen cache of packed permeation, the
compartment and unwrapped position of packed myelinated cylinders and the echo
counter of CPMG.  The refusals (a split walk on a plan without the parked phase,
another seed or another output set mid-walk) are checked on the host path of the
same calls.
"""
import numpy as np
import pytest

import dmipy_sim_amd as ds
from substrates import D, packed_cyl, packed_sph, small_pgse

pytestmark = pytest.mark.gpu

PARKED = ds._native.PHASE_FORMS[3]

OUTS = ("final_pos", "comp_origin", "comp_final", "comp_full", "pos_full", "n_cross", "log_w",
        "phi", "signal_im")


def _pmyel():
    rng = np.random.default_rng(0)
    ri, gr, c = ds.pack_myelinated_cylinders(rng.uniform(1e-6, 2e-6, 8), 0.7, 0.5, seed=1)
    L = float(np.sqrt(np.pi * np.sum((ri / gr) ** 2) / 0.5))
    return ds.PackedMyelinatedCylinders(ri, gr, c, L, N_max=16, D_myelin=5e-10, kappa_inner=3e-5,
                                        kappa_outer=2e-5, rho_outer=2e-5, rho_inner=1e-5,
                                        T2_myelin=0.02)


# (name, geometry, T2, slices): n_sub >= 32 and 20 measurements -> parked phase
CASES = [
    ("pcyl_perm", lambda: packed_cyl(kappa=5e-5, rho=1e-5), None, (7, 1, 13)),
    ("psph_perm", lambda: packed_sph(kappa=5e-5), None, (15,)),
    # 24 spheres at VF 0.5: grid cells with shrunk boxes (levels 1 and 2)
    ("psph_dense", lambda: packed_sph(kappa=5e-5, rho=1e-5, n=24, vf=0.5, seed=6), None, (9, 4, 17)),
    ("cyl_perm", lambda: ds.Cylinder(3e-6, (0, 0, 1), permeability=5e-5), None, (1, 1, 1, 10)),
    ("sph_rho", lambda: ds.Sphere(3e-6, surface_relaxivity_t2=2e-5), 0.06, (11, 11)),
    ("pmyel", _pmyel, None, (4, 20)),
]


def _same(got, ref, keys):
    np.testing.assert_array_equal(got["signal"], ref["signal"])
    for k in keys:
        np.testing.assert_array_equal(got[k], ref[k], err_msg=k)
    assert got["n_hits"] == ref["n_hits"] and got["n_cand"] == ref["n_cand"]


@pytest.mark.parametrize("name,factory,T2,slices", CASES, ids=[c[0] for c in CASES])
def test_sliced_walk_equals_one_call(gpu, name, factory, T2, slices):
    wf = small_pgse(n_dirs=5)
    geom = factory()
    print(f"resume {name}: 2000 walkers, 20 measurements x 30 samples, slices {slices}", flush=True)
    ref = gpu.walk(geom, D, wf, 2000, seed=21, T2=T2, outputs=OUTS, counters=True)
    assert ref["phase_form"] == PARKED and ref["n_sub"] >= 32
    got = gpu.walk(geom, D, wf, 2000, seed=21, T2=T2, outputs=OUTS, counters=True, slices=slices)
    _same(got, ref, OUTS)
    assert got["n_hits"] > 0


def test_sliced_cpmg_echo_rows(gpu):
    wf = small_pgse(n_dirs=5)
    geom = ds.Sphere(3e-6, surface_relaxivity_t2=2e-5)
    echo = np.array([4, 9, 10, 29], np.int32)
    keys = ("log_w", "signal_im", "final_pos")
    print("resume cpmg: sphere, echoes at samples 4 9 10 29, slices 4 5 1 3", flush=True)
    ref = gpu.walk(geom, D, wf, 2000, seed=5, T2=0.05, outputs=keys, echo_indices=echo, counters=True)
    # boundaries at 4, 9, 10, 13: before, on and between echo samples
    got = gpu.walk(geom, D, wf, 2000, seed=5, T2=0.05, outputs=keys, echo_indices=echo, counters=True,
                   slices=(4, 5, 1, 3))
    assert ref["signal"].shape == (4, 20)
    _same(got, ref, keys)


def test_walker_shard_resumes(gpu):
    # a shard [first_walker, first_walker + N) keeps its own streams across slices
    wf = small_pgse(n_dirs=5)
    geom = packed_cyl(kappa=5e-5)
    print("resume shard: packed cylinders, walkers [3000, 4500) of 6000", flush=True)
    kw = dict(seed=3, outputs=("final_pos", "n_cross"), first_walker=3000, n_total=6000)
    ref = gpu.walk(geom, D, wf, 1500, **kw)
    got = gpu.walk(geom, D, wf, 1500, slices=(2, 2, 2), **kw)
    np.testing.assert_array_equal(got["signal"], ref["signal"])
    np.testing.assert_array_equal(got["final_pos"], ref["final_pos"])
    np.testing.assert_array_equal(got["n_cross"], ref["n_cross"])


def test_planned_walk_advance_and_refusals(gpu):
    wf = small_pgse(n_dirs=5)
    geom = packed_cyl(kappa=5e-5)
    print("resume plan: advance in 3 calls, rewind, refusals", flush=True)
    pw = gpu.PlannedWalk(geom, D, wf, 1200)
    try:
        ref = pw.run(seed=11, counters=True)
        assert ref["phase_form"] == PARKED
        steps = [pw.advance(10, seed=11, counters=True) for _ in range(3)]
        assert [s["t_done"] for s in steps] == [10, 20, 30]
        assert "signal" not in steps[0] and steps[0]["kernel_ms"] > 0
        np.testing.assert_array_equal(steps[-1]["signal"], ref["signal"])
        assert steps[-1]["n_hits"] == ref["n_hits"]
        # mid-walk: another seed, another output set
        assert pw.advance(5, seed=11)["t_done"] == 5
        with pytest.raises(ValueError, match="seed"):
            pw.advance(5, seed=12)
        with pytest.raises(ValueError, match="outputs"):
            pw.advance(5, seed=11, counters=True)
        # rewind abandons the walk: a new one starts from sample 0 with any seed
        pw.rewind()
        a = pw.advance(30, seed=11, counters=True)
        np.testing.assert_array_equal(a["signal"], ref["signal"])
        with pytest.raises(ValueError, match="n_samples"):
            pw.advance(0, seed=11)
    finally:
        pw.close()
    # a plan without the parked phase walks only in one call
    cyl = ds.Cylinder(3e-6, (0, 0, 1))
    pw = gpu.PlannedWalk(cyl, D, wf, 800)
    try:
        one = pw.run(seed=4)
        assert one["phase_form"] != PARKED
        with pytest.raises(NotImplementedError, match="parked"):
            pw.advance(10, seed=4)
        whole = pw.advance(30, seed=4)
        np.testing.assert_array_equal(whole["signal"], one["signal"])
    finally:
        pw.close()
