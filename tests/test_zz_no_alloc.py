"""dmw_walk_noalloc: the walk with every device buffer in the library's static
scratch array (no device allocation) equals dmw_walk on the same inputs.

Two calls in a row through the C ABI, the second smaller (fewer walkers, fewer
measurements, another geometry), each compared with dmw_walk: the same kernels
in the same order, so signals and per-walker outputs are identical.  A walk whose
buffers exceed the array is refused on the host before any launch.
"""
import numpy as np
import pytest

import dmipy_sim_amd as ds
from substrates import D, packed_cyl, small_pgse

pytestmark = pytest.mark.gpu

OUTS = ("final_pos", "comp_origin", "comp_final", "n_cross", "log_w", "phi", "signal_im")


def _check_same(a, b):
    np.testing.assert_array_equal(a["signal"], b["signal"])
    for k in OUTS:
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)


def test_noalloc_walk_equals_walk_twice(gpu):
    wf = small_pgse()
    geom = packed_cyl(kappa=5e-5, rho=1e-5)
    print("noalloc: call 1, packed cylinders, 3000 walkers, 16 measurements", flush=True)
    ref = gpu.walk(geom, D, wf, 3000, seed=8, outputs=OUTS)
    got = gpu.walk(geom, D, wf, 3000, seed=8, outputs=OUTS, noalloc=True)
    _check_same(got, ref)
    assert got["n_cross"].sum() > 0
    wf2 = small_pgse(n_dirs=2, b=(1e9, 2e9))
    geom2 = ds.Cylinder(3e-6, (0, 0, 1), surface_relaxivity_t2=2e-5)
    print("noalloc: call 2, cylinder, 700 walkers, 4 measurements", flush=True)
    ref2 = gpu.walk(geom2, D, wf2, 700, seed=9, outputs=OUTS)
    got2 = gpu.walk(geom2, D, wf2, 700, seed=9, outputs=OUTS, noalloc=True)
    _check_same(got2, ref2)


def test_noalloc_rejects_walks_beyond_the_scratch_array(gpu):
    wf = small_pgse()
    geom = ds.Sphere(3e-6)
    # phi alone: 16 measurements x 4 bytes x 1.2e6 walkers = 77 MB > 64 MiB
    print("noalloc: oversize request (1.2e6 walkers with per-walker phases)", flush=True)
    with pytest.raises(ValueError, match="scratch array"):
        gpu.walk(geom, D, wf, 1_200_000, seed=1, outputs=("phi",), noalloc=True)
    # the array is still usable afterwards
    ref = gpu.walk(geom, D, wf, 500, seed=2, outputs=OUTS)
    got = gpu.walk(geom, D, wf, 500, seed=2, outputs=OUTS, noalloc=True)
    _check_same(got, ref)
