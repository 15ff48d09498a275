"""The walk's wall screen never clears a sub-step the reference would stop.

The packed-substrate walk skips the reference's per-object root test
(geometries.py:2067-2093, 2667-2691) for every sub-step that its screen
(dmipy-sim_amd/csrc/screen.h) clears.  That is exact only if a cleared sub-step
has no valid root in the reference's own float32 arithmetic.  Here the screen,
compiled for the host exactly as the kernel evaluates it (the cached image centre
c + L k, f = |x - c|^2 - R^2 at both ends of the step), is held against the
oracle's float32 root test on millions of adversarial configurations: walkers
within a few steps of the wall on both sides, near-tangent steps, grazing steps
that end exactly on the wall, radii from 50 nm to 12 um, steps from R/400 to R/6,
coordinates out to the edge of the periodic cell.  A cleared sub-step with a
valid root is a failure; the screen must also clear every step that cannot reach
the wall, so it is not vacuous.
"""
import ctypes as C
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SCREEN_LIB = os.path.join(ROOT, "tests", "gpu_selftest", "libscreen_check.so")

_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int32)


@pytest.fixture(scope="module")
def screen():
    if not os.path.exists(SCREEN_LIB):
        import subprocess
        subprocess.run(["make", "-C", os.path.dirname(SCREEN_LIB), "libscreen_check.so"], check=True)
    h = C.CDLL(SCREEN_LIB)
    h.screen_check_clear.restype = None
    h.screen_check_clear.argtypes = [C.c_int64, _fp, _fp, _fp, _fp, C.c_float, C.c_float,
                                     C.c_int32, _ip]
    h.screen_check_margin.restype = C.c_double
    h.screen_check_margin.argtypes = [C.c_double, C.c_double, C.c_double]
    h.screen_check_grid.restype = C.c_int64
    h.screen_check_grid.argtypes = [C.c_int32, _fp, _fp, C.c_float, C.c_float, C.c_float, C.c_int32,
                                    C.c_double, C.c_int32, C.c_int32, C.c_uint64, C.POINTER(C.c_int64),
                                    C.POINTER(C.c_int64)]
    return h


def _roots(oracle, p, s, c, R, L, eps, dim):
    lib = oracle.lib()
    fn = lib.oracle_permeate_roots
    fn.restype = None
    fn.argtypes = [C.c_int64, _fp, _fp, _fp, _fp, C.c_float, C.c_float, C.c_int32, _ip]
    out = np.zeros(len(R), np.int32)
    fn(len(R), p.ctypes.data_as(_fp), s.ctypes.data_as(_fp), c.ctypes.data_as(_fp),
       R.ctypes.data_as(_fp), C.c_float(L), C.c_float(eps), dim, out.ctypes.data_as(_ip))
    return out


def _clear(screen, p, s, cc, R, margin, step, dim):
    out = np.zeros(len(R), np.int32)
    screen.screen_check_clear(len(R), p.ctypes.data_as(_fp), s.ctypes.data_as(_fp),
                              cc.ctypes.data_as(_fp), R.ctypes.data_as(_fp), C.c_float(margin),
                              C.c_float(step), dim, out.ctypes.data_as(_ip))
    return out


def _cases(rng, n, dim, L, R, step):
    """Walkers near the wall of an object of radius R, with 3-D steps of length
    `step` whose first `dim` components are the geometry's.  Walkers sit anywhere
    in the frame the kernel walks in (the cell plus a sample's drift past its
    edge); the object's centre is the image at the chosen offset, folded back into
    the cell as the substrate stores it."""
    f32 = np.float32
    u = rng.standard_normal((n, dim))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    kind = rng.integers(0, 4, n)
    d = np.where(kind == 0, rng.uniform(-3, 3, n) * step,                # near the wall
        np.where(kind == 1, rng.uniform(-1.2, 1.2, n) * step,           # within reach
        np.where(kind == 2, rng.standard_normal(n) * 1e-3 * step,       # on the wall
                 rng.uniform(0.8, 1.2, n) * step * np.sign(rng.standard_normal(n)))))
    p = np.zeros((n, 3), np.float64)
    p[:, :dim] = rng.uniform(-0.5 * L - 3 * step, 0.5 * L + 3 * step, (n, dim))
    p = p.astype(f32).astype(np.float64)
    cc = p[:, :dim] - (R + d)[:, None] * u
    c = np.zeros((n, 3), f32)
    c[:, :dim] = cc - L * np.floor(cc / L + 0.5)
    # steps: isotropic, a quarter of them near tangent to the wall
    s = rng.standard_normal((n, 3))
    tang = rng.random(n) < 0.25
    if dim == 2:
        perp = np.stack([-u[:, 1], u[:, 0]], 1)
        s[tang, :2] = perp[tang] + rng.standard_normal((tang.sum(), 1)) * 1e-3 * u[tang]
        s[tang, 2] = rng.standard_normal(tang.sum()) * rng.random(tang.sum())
    else:
        t3 = np.cross(u, rng.standard_normal((n, 3)))
        s[tang] = t3[tang] + rng.standard_normal((tang.sum(), 1)) * 1e-3 * u[tang]
    s /= np.linalg.norm(s, axis=1, keepdims=True)
    s *= step
    return (np.ascontiguousarray(p, f32), np.ascontiguousarray(s, f32), np.ascontiguousarray(c, f32))


def _image_centre(p, c, L, dim):
    """The kernel's cached image centre: c + L k with k = floor((p - c)/L + 0.5)."""
    f32 = np.float32
    cc = c.copy()
    inv = f32(1.0) / f32(L)
    for k in range(dim):
        q = (p[:, k] - c[:, k]).astype(f32)
        sh = np.floor(q * inv + f32(0.5)).astype(f32)
        cc[:, k] = (c[:, k] + f32(L) * sh).astype(f32)
    return cc


@pytest.mark.parametrize("dim", [2, 3])
def test_screen_never_clears_a_reference_hit(screen, oracle, dim):
    rng = np.random.default_rng(100 + dim)
    checked = hits = flagged = 0
    for L in (58e-6, 200e-6):
        for R in (5e-8, 3e-7, 5.8e-7, 3e-6, 1.2e-5):
            for frac in (1 / 6, 1 / 25, 1 / 400):
                step = R * frac
                if R > 0.4 * L:
                    continue
                p, s, c = _cases(rng, 200_000, dim, L, R, step)
                Rv = np.full(len(p), R, np.float32)
                cc = _image_centre(p, c, L, dim)
                valid = _roots(oracle, p, s, c, Rv, L, 1e-7 * R, dim)
                m = screen.screen_check_margin(L, R, step)
                for scale in (1.0, 0.25):          # the bound holds with room to spare
                    clear = _clear(screen, p, s, cc, Rv, m * scale, step, dim)
                    bad = np.flatnonzero((clear == 1) & (valid == 1))
                    assert bad.size == 0, (L, R, step, scale, p[bad[0]], s[bad[0]], c[bad[0]])
                clear = _clear(screen, p, s, cc, Rv, m, step, dim)
                # not vacuous: steps whose end stays farther than one step plus the
                # margins from the wall are cleared
                gap = np.abs(np.linalg.norm((p - cc)[:, :dim].astype(np.float64), axis=1) - R)
                far = gap > step + 2 * m + step * step / R
                assert np.all(clear[far] == 1)
                checked += len(p)
                hits += int(valid.sum())
                flagged += int((clear == 0).sum())
    assert hits > 10_000 and checked > 4_000_000
    # flags are mostly real hits (the near-wall sample is dense in hits by design)
    assert flagged < 3 * hits + 0.05 * checked


def test_margin_is_small_against_the_step(screen):
    """On the packed-cylinder workload (cell 58 um, radii up to ~6 um, steps of
    23 nm) the margin is a few percent of a step."""
    m = screen.screen_check_margin(58e-6, 6.2e-6, 2.32e-8)
    assert 5e-10 < m < 0.05 * 2.32e-8


@pytest.mark.parametrize("which", ["packed cylinders, 64, VF 0.6", "small cylinders", "small spheres"])
@pytest.mark.parametrize("max_list", [0, 3])
def test_expanded_cells_list_every_reachable_object(screen, which, max_list):
    """The screened path keeps a walker on its cell's list while it stays inside
    the cell box expanded by half a side (less for a cell whose list would not fit
    the walker's object cache of three, at the box level the list start carries):
    every point of that box within one step of an object's wall must find the
    object in the list (the builder inflates the box per coordinate, as the walk
    tests it)."""
    import dmipy_sim_amd as ds
    from dmipy_sim_amd import configs
    from substrates import D, packed_cyl, packed_sph, small_pgse
    if which.startswith("packed"):
        cfg = configs.packed_cylinders_free_waveform()
        g, dt, Dw, dim = cfg["geometry"], cfg["waveform"].dt, cfg["diffusivity"], 2
    else:
        g = packed_cyl(kappa=5e-5) if "cyl" in which else packed_sph(kappa=5e-5)
        dt, Dw, dim = small_pgse().dt, D, 2 if "cyl" in which else 3
    plan = ds.make_walk_plan(g, Dw, dt)
    gd = plan.geometry_desc()
    c = np.ascontiguousarray(g._centers_np, np.float32)
    r = np.ascontiguousarray(g._radii_np, np.float32)
    pairs = C.c_int64(0)
    stats = (C.c_int64 * 5)()
    # a probe step of a quarter cell side: many near-wall pairs per cell
    for step in (plan.step_len, 0.05 * float(g._L_float)):
        missing = screen.screen_check_grid(len(r), c.ctypes.data_as(_fp), r.ctypes.data_as(_fp),
                                           float(np.float32(g._L_float)), step, gd.nudge, dim, 0.5,
                                           max_list, 40 if dim == 3 else 200, 7, C.byref(pairs), stats)
        assert pairs.value > 1000
        assert missing == 0
        if max_list == 0:
            assert stats[1] == 0 and stats[2] == 0
        if which.startswith("packed") and max_list and step == plan.step_len:
            # the 64-cylinder packing at its own step: a few cells of 4 objects at the
            # full box, none after the builder's shrinking, most cells at the full box
            assert stats[1] + stats[2] >= 1 and stats[4] == 0 and stats[3] <= 3
            assert stats[1] + stats[2] < 0.01 * stats[0]


def test_box_levels_shrink_only_overflowing_cells(screen):
    """Against the plain grid (no cache limit), the limited grid changes only the
    lists that exceed three objects: those get a smaller box and a shorter list,
    a subset of the full-box list; every other cell keeps its list and level 0."""
    from dmipy_sim_amd import configs
    cfg = configs.packed_cylinders_free_waveform()
    g = cfg["geometry"]
    import dmipy_sim_amd as ds
    plan = ds.make_walk_plan(g, cfg["diffusivity"], cfg["waveform"].dt)
    c = np.ascontiguousarray(g._centers_np, np.float32)
    r = np.ascontiguousarray(g._radii_np, np.float32)
    fn = screen.screen_check_lists
    fn.restype = C.c_int64
    fn.argtypes = [C.c_int32, _fp, _fp, C.c_float, C.c_float, C.c_float, C.c_int32, C.c_double,
                   C.c_int32, _ip, _ip, _ip, C.c_int64]
    cap = 1 << 20
    out = [np.zeros(cap, np.int32) for _ in range(3)]
    lens = {}
    for max_list in (0, 3):
        o = [np.zeros(cap, np.int32) for _ in range(3)]
        n_cells = fn(len(r), c.ctypes.data_as(_fp), r.ctypes.data_as(_fp), float(np.float32(g._L_float)),
                     plan.step_len, plan.geometry_desc().nudge, 2, 0.5, max_list,
                     o[0].ctypes.data_as(_ip), o[1].ctypes.data_as(_ip), o[2].ctypes.data_as(_ip), cap)
        assert n_cells > 0
        start, level, items = o[0][:n_cells + 1], o[1][:n_cells], o[2][:o[0][n_cells]]
        lens[max_list] = (start, level, items)
    s0, l0, i0 = lens[0]
    s3, l3, i3 = lens[3]
    assert not l0.any()
    changed = 0
    for cell in range(len(l3)):
        a = list(i0[s0[cell]:s0[cell + 1]])
        b = list(i3[s3[cell]:s3[cell + 1]])
        if len(a) <= 3:
            assert l3[cell] == 0 and a == b
        else:
            changed += 1
            assert l3[cell] in (1, 2) and len(b) <= 3 and set(b) <= set(a) and b == sorted(b)
    assert changed >= 1
