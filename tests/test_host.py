"""Host-side logic: sub-step rules and float32 constants, packers, the C ABI
surface, and argument handling of simulate() that runs before any GPU work."""
import os
import re

import numpy as np
import pytest

import dmipy_sim_amd as ds
from dmipy_sim_amd import _native
from substrates import CASES, WALK_CASES, D

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("name,factory,T2,T1", WALK_CASES, ids=[c[0] for c in WALK_CASES])
@pytest.mark.parametrize("dt", [1e-4, 2.5e-5, 5.005e-5])
def test_plan_constants_match_oracle_restatement(oracle, name, factory, T2, T1, dt):
    """The product's plan and the oracle's independent restatement of
    make_step_fn agree on n_sub and every float32 constant, bit for bit."""
    g = factory()
    p = ds.make_walk_plan(g, D, dt, T2=T2, T1=T1)
    o = oracle.oracle_plan(g, D, dt, T2, T1)
    assert p.n_sub == o["n_sub"] and p.wall == o["wall"] and p.kind == o["kind"]
    for k in ("step_len", "gamma_dt_sub", "dt_sub", "rho_over_D", "kappa_over_D",
              "inv_T2", "inv_T1", "eps", "nudge", "size", "cell", "aux"):
        assert np.float32(getattr(p, k)) == np.float32(o[k]), k
    assert [np.float32(v) for v in p.axes] == [np.float32(v) for v in o["axes"]]
    if p.myelin is not None:
        for pk, ok in (("step_len", "m_step"), ("two_kappa_over_D", "m_k2"),
                       ("dt_over_T2", "m_q2"), ("spin_weight", "m_wf")):
            assert [np.float32(v) for v in p.myelin[pk]] == [np.float32(v) for v in o[ok]], pk
        for N in (1, 7, 1000, 12345):
            n0, n1, _ = p.myelin["start_counts"](N)
            assert (n0, n1) == oracle.myelin_start_counts(o, N)


def test_substep_rules_closed_form():
    """R/25 permeable (dt_max = R^2/(3750 D)), R/8 surface, 1 otherwise
    (physics.py:35-48, 163-188); Cylinder PGSE workload gives n_sub = 2."""
    dt = 0.05 / 999
    assert ds.make_walk_plan(ds.Cylinder(5e-6, (0, 0, 1), surface_relaxivity_t2=1e-6), D, dt).n_sub == 2
    assert ds.make_walk_plan(ds.Cylinder(5e-6, (0, 0, 1)), D, dt).n_sub == 1
    R = 3e-6
    assert ds.make_walk_plan(ds.Sphere(R, permeability=1e-5), D, dt).n_sub == \
        int(np.ceil(dt / (R * R / (3750 * D))))
    assert ds.make_walk_plan(ds.Box1D(5e-6, surface_relaxivity_t2=1e-5), D, dt).n_sub == 1
    g = ds.Sphere(R, surface_relaxivity_t2=1e-5)
    g.surface_substep_frac = 0
    assert ds.make_walk_plan(g, D, dt).n_sub == 1
    assert ds.make_walk_plan(ds.Sphere(R, surface_relaxivity_t2=1e-5), D, dt,
                             substep_rule="replay").n_sub == int(np.ceil(dt / (R * R / (216 * D))))


def test_packed_workload_substeps():
    """64 Gamma(4, 0.75 um) radii, seed 0: smallest radius 0.58 um -> n_sub = 558."""
    from dmipy_sim_amd import configs
    cfg = configs.packed_cylinders_free_waveform()
    g = cfg["geometry"]
    assert len(g._radii_np) == 64 and abs(g.volume_fraction() - 0.6) < 1e-12
    assert ds.make_walk_plan(g, D, cfg["waveform"].dt).n_sub == 558
    assert g.min_gap > 0


# ---------------------------------------------------------------- packers
def _no_overlap(c, r, L):
    for i in range(len(r)):
        for j in range(i + 1, len(r)):
            dq = c[i] - c[j]
            dq -= L * np.round(dq / L)
            assert np.linalg.norm(dq) >= r[i] + r[j] - 1e-10


def test_pack_cylinders_rsa_invariants():
    """tests/test_packed_cylinders.py:27-76."""
    c, L, vf = ds.pack_cylinders(np.full(8, 2e-6), target_vf=0.20, seed=42)
    assert c.shape == (8, 2)
    _no_overlap(c, np.full(8, 2e-6), L)
    radii = np.random.default_rng(0).uniform(1e-6, 4e-6, 6)
    c, L, vf = ds.pack_cylinders(radii, target_vf=0.15, seed=7)
    _no_overlap(c, radii, L)
    c, L, vf = ds.pack_cylinders(np.full(5, 3e-6), target_vf=0.20, seed=1)
    assert vf == pytest.approx(np.pi * 5 * 9e-12 / L ** 2, rel=1e-6)
    assert vf == pytest.approx(0.20, rel=1e-3)


def test_pack_errors():
    with pytest.raises(ValueError, match="exactly one"):
        ds.pack_cylinders([1e-6], target_vf=0.2, L=100e-6)
    with pytest.raises(ValueError, match="exactly one"):
        ds.pack_cylinders([1e-6])
    with pytest.raises(ValueError, match="positive"):
        ds.pack_cylinders([-1e-6], target_vf=0.1)
    with pytest.raises(ValueError, match="target_vf"):
        ds.pack_cylinders([1e-6], target_vf=1.5)
    with pytest.raises(ValueError, match="exceeds"):
        ds.pack_spheres([30e-6], L=40e-6)


def test_pack_spheres_and_dense_packing():
    r = np.full(10, 2e-6)
    c, L, vf = ds.pack_spheres(r, target_vf=0.2, seed=3)
    _no_overlap(c, r, L)
    radii = np.random.default_rng(5).gamma(4.0, 0.75e-6, 64)
    c, L, vf = ds.pack_dense(radii, 0.6, seed=0, dim=2, gap=0.1)
    assert vf == pytest.approx(0.6, rel=1e-12)
    _no_overlap(c, radii, L)
    g = ds.PackedCylinders(radii, c, L)
    assert g.min_gap >= 0.1 * radii.min() * 0.999


def test_packed_init_positions_outside():
    radii = np.full(6, 2e-6)
    c, L, _ = ds.pack_spheres(radii, target_vf=0.2, seed=1)
    g = ds.PackedSpheres(radii, c, L)
    p = g.init_positions(2000, 3)
    for ci, ri in zip(c, radii):
        q = p - ci
        q -= L * np.round(q / L)
        assert np.all(np.linalg.norm(q, axis=1) > ri)


def test_geometry_closed_forms():
    """tests/test_geometry_properties.py: volumes, areas, volume fractions."""
    s = ds.Sphere(2e-6)
    assert s.volume() == pytest.approx(4 / 3 * np.pi * 8e-18)
    assert s.surface_area() == pytest.approx(4 * np.pi * 4e-12)
    cy = ds.Cylinder(2e-6, (1, 1, 0))
    assert cy.volume() == pytest.approx(np.pi * 4e-12)
    assert cy.surface_area(include_caps=True) == pytest.approx(2 * np.pi * 2e-6 + 2 * np.pi * 4e-12)
    np.testing.assert_allclose(cy._R.astype(np.float64) @ cy.orientation, [0, 0, 1], atol=1e-6)


# ---------------------------------------------------------------- C ABI
def _header_symbols():
    txt = open(os.path.join(ROOT, "include", "dmipy_walk.h")).read()
    return set(re.findall(r"^\s*(?:int|int64_t|const char\*)\s+(dmw_\w+)\s*\(", txt, re.M))


def test_library_exports_every_declared_symbol():
    lib = _native.lib()
    decl = _header_symbols()
    assert decl == set(_native.SIGNATURES), decl ^ set(_native.SIGNATURES)
    for name in decl:
        assert hasattr(lib, name), name


def test_struct_layouts_match_header(tmp_path):
    """Every field of the ctypes mirrors sits at the offset gcc gives it when
    compiling include/dmipy_walk.h, and every struct has gcc's size."""
    import ctypes as C
    import subprocess
    structs = {"dmw_geometry": _native.Geometry_t, "dmw_waveform": _native.Waveform_t,
               "dmw_run": _native.Run_t, "dmw_result": _native.Result_t,
               "dmw_instance": _native.Instance_t, "dmw_batch_stats": _native.BatchStats_t,
               "dmw_replay": _native.Replay_t}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "dmipy_walk.h"',
             'int main(void) {']
    for cname, ct in structs.items():
        lines.append(f'printf("{cname} __sizeof__ %zu\\n", sizeof({cname}));')
        for fname, _ in ct._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    seen = 0
    for line in filter(None, out):
        cname, field, val = line.split()
        ct = structs[cname]
        got = C.sizeof(ct) if field == "__sizeof__" else getattr(ct, field).offset
        assert got == int(val), (cname, field, got, val)
        seen += 1
    assert seen == sum(len(ct._fields_) + 1 for ct in structs.values())


def test_device_count_without_gpu_is_zero_or_more():
    assert _native.device_count() >= 0


def test_simulate_argument_errors_before_gpu():
    wf = ds.pgse(0.01, 0.04, 0.05, [[1, 0, 0]], 50)
    with pytest.raises(ValueError, match="return_compartments"):
        ds.simulate(10, D, wf, ds.FreeDiffusion(), return_compartments="bogus")
    with pytest.raises(ValueError, match="engine"):
        ds.simulate(10, D, wf, ds.FreeDiffusion(), engine="jax")
    with pytest.raises(NotImplementedError, match="replay"):
        ds.simulate(10, D, wf, ds.Sphere(3e-6, permeability=1e-5), engine="replay")
    with pytest.raises(ValueError, match="echo_indices"):
        ds.simulate_cpmg(10, D, wf, ds.FreeDiffusion())


def test_no_cpu_fallback_without_gpu():
    """Without a device the walk raises; there is no CPU path to fall back to."""
    if _native.device_count() > 0:
        pytest.skip("a GPU is visible")
    wf = ds.pgse(0.01, 0.04, 0.05, [[1, 0, 0]], 50)
    with pytest.raises(RuntimeError, match="no HIP device"):
        ds.simulate(10, D, wf, ds.FreeDiffusion(), require_gpu=False)
    with pytest.raises(RuntimeError):
        ds.simulate(10, D, wf, ds.FreeDiffusion(), require_gpu=True)


def test_oom_backoff_halves_walker_batches(monkeypatch):
    """simulate() retries at half the walker batch on "out of memory"
    (core.py:366-394) and recombines chunks with seeds seed+1+b (core.py:765-818).
    The walk is replaced by a host stub that fails above 250 walkers."""
    from dmipy_sim_amd import core
    calls = []

    def fake_walk(geometry, diffusivity, waveform, n_walkers, seed=123, **kw):
        calls.append((n_walkers, seed))
        if n_walkers > 250:
            raise RuntimeError("hipMalloc: out of memory")
        return {"signal": np.full(waveform.G.shape[0], float(seed)), "kernel_ms": 0.0}

    monkeypatch.setattr(core, "walk", fake_walk)
    monkeypatch.setattr(core, "check_gpu", lambda *a, **k: True)
    wf = ds.pgse(0.01, 0.04, 0.05, [[1, 0, 0]], 20)
    with pytest.warns(RuntimeWarning, match="OOM"):
        s = ds.simulate(1000, D, wf, ds.FreeDiffusion(), seed=10)
    sizes = [c[0] for c in calls]
    assert sizes[0] == 1000 and 250 in sizes
    ok = [c for c in calls if c[0] <= 250]
    assert [c[1] for c in ok] == [11, 12, 13, 14]
    assert s[0] == pytest.approx(np.mean([11, 12, 13, 14]))


def test_replay_routes_before_walker_batches(monkeypatch):
    """engine='replay' is routed before the fused engine's OOM back-off and
    walker batching (core.py:334-349): one producer walk under the caller's
    seed, whatever walker_batch_size says; AcquisitionScheme-style wrappers are
    unwrapped.  The replay and the fused walk are host stubs."""
    from dmipy_sim_amd import core
    calls = []

    def fake_replay(n_walkers, diffusivity, waveform, geometry, *, seed, T2, T1, r0):
        calls.append((n_walkers, seed, waveform))
        return np.full(waveform.G.shape[0], 0.5, np.float32)

    def no_walk(*a, **k):
        raise AssertionError("the fused walk must not run for engine='replay'")

    monkeypatch.setattr(core, "_simulate_via_replay", fake_replay)
    monkeypatch.setattr(core, "walk", no_walk)
    monkeypatch.setattr(core, "check_gpu", lambda *a, **k: True)
    wf = ds.pgse(0.01, 0.04, 0.05, [[1, 0, 0]], 20)

    class Scheme:
        waveform = wf

    s = ds.simulate(1000, D, Scheme(), ds.Cylinder(5e-6, (0, 0, 1)), seed=10, walker_batch_size=100,
                    engine="replay")
    assert calls == [(1000, 10, wf)]
    assert s.dtype == np.float32 and s.shape == (1,)


def test_batch_argument_errors_before_gpu():
    """The batched launch validates its descriptors before touching a device."""
    import ctypes as C
    wf = ds.pgse(0.01, 0.04, 0.05, [[1, 0, 0]], 20)
    with pytest.raises(ValueError, match="64"):
        ds.BatchedWalk([dict(geometry=ds.Sphere(3e-6), diffusivity=D, n_walkers=64)],
                       ds.tile_waveform(wf, 65))
    with pytest.raises(ValueError, match="waveform index"):
        ds.BatchedWalk([dict(geometry=ds.Sphere(3e-6), diffusivity=D, n_walkers=64,
                             waveform=1)], wf)
    with pytest.raises(NotImplementedError, match="T1"):
        ds.BatchedWalk([dict(geometry=ds.Sphere(3e-6), diffusivity=D, n_walkers=64,
                             T1=1.0)], wf)
    G = np.ascontiguousarray(wf.G, np.float32)
    wd = (_native.Waveform_t * 1)()
    wd[0] = _native.Waveform_t(n_meas=1, n_t=G.shape[1], G=_native.fptr(G), chi=None,
                               n_echo=0, echo_idx=None)
    inst = (_native.Instance_t * 1)()
    gd = _native.Geometry_t(kind=3, wall=0, n_obj=0, size=3e-6, cell=0.0, eps=3e-13,
                            nudge=3e-10, rho_over_D=0.0, kappa_over_D=0.0,
                            centers=None, radii=None)
    h = C.c_void_p()
    lib = _native.lib()
    for bad in (dict(wave_set=1), dict(n_walkers=0), dict(n_sub=0)):
        kw = dict(geom=gd, n_walkers=64, n_sub=1, step_len=1e-7, gamma_dt_sub=1.0,
                  dt_sub=1e-5, inv_T2=0.0, has_t2=0, wave_set=0)
        kw.update(bad)
        inst[0] = _native.Instance_t(**kw)
        assert lib.dmw_batch_create(C.byref(h), 1, inst, 1, wd, None, 1) == _native.DMW_EINVAL
    gd.kind = 4       # packed without objects
    inst[0] = _native.Instance_t(geom=gd, n_walkers=64, n_sub=1, step_len=1e-7,
                                 gamma_dt_sub=1.0, dt_sub=1e-5, inv_T2=0.0, has_t2=0, wave_set=0)
    assert lib.dmw_batch_create(C.byref(h), 1, inst, 1, wd, None, 1) == _native.DMW_EINVAL
    gd.kind = 0
    gd.wall = 1       # free diffusion has no wall to relax on
    inst[0] = _native.Instance_t(geom=gd, n_walkers=64, n_sub=1, step_len=1e-7,
                                 gamma_dt_sub=1.0, dt_sub=1e-5, inv_T2=0.0, has_t2=0, wave_set=0)
    assert lib.dmw_batch_create(C.byref(h), 1, inst, 1, wd, None, 1) == _native.DMW_EUNSUPPORTED


def test_device_arch_arguments():
    """dmw_device_arch refuses a null or empty buffer before touching a device,
    and without a device reports an error rather than a name."""
    import ctypes as C
    lib = _native.lib()
    assert lib.dmw_device_arch(0, None, 16) == _native.DMW_EINVAL
    buf = C.create_string_buffer(8)
    assert lib.dmw_device_arch(0, buf, 0) == _native.DMW_EINVAL
    if _native.device_count() == 0:
        assert lib.dmw_device_arch(0, buf, len(buf)) != _native.DMW_OK
        assert _native.device_arch() is None


def test_batch_tuning_checked_before_gpu():
    """dmw_batch_create refuses launch shapes outside the compiled set before
    touching a device."""
    import ctypes as C
    wf = ds.pgse(0.01, 0.04, 0.05, [[1, 0, 0]], 20)
    G = np.ascontiguousarray(wf.G, np.float32)
    wd = (_native.Waveform_t * 1)()
    wd[0] = _native.Waveform_t(n_meas=1, n_t=20, G=_native.fptr(G), chi=None, n_echo=0, echo_idx=None)
    gd = _native.Geometry_t(kind=0, wall=0, size=0.0, eps=0.0, nudge=0.0, rho_over_D=0.0,
                            kappa_over_D=0.0, n_obj=0, cell=0.0, centers=None, radii=None)
    inst = (_native.Instance_t * 1)()
    inst[0] = _native.Instance_t(geom=gd, n_walkers=64, n_sub=1, step_len=1e-7, gamma_dt_sub=1.0,
                                 dt_sub=1e-5, inv_T2=0.0, has_t2=0, wave_set=0)
    lib = _native.lib()
    h = C.c_void_p()
    for bad, text in ((dict(walkers_per_wave=100), "walkers_per_wave"),
                      (dict(walkers_per_wave=1024), "walkers_per_wave"),
                      (dict(unroll=3), "unroll"), (dict(waves_per_simd=9), "waves_per_simd"),
                      (dict(waves_per_simd=-1), "waves_per_simd")):
        tn = _native.BatchTuning_t(**bad)
        assert lib.dmw_batch_create(C.byref(h), 1, inst, 1, wd, C.byref(tn), 1) == _native.DMW_EINVAL
        assert text in lib.dmw_last_error().decode()
    with pytest.raises(ValueError, match="tuning"):
        ds.BatchedWalk([dict(geometry=ds.Sphere(3e-6), diffusivity=D, n_walkers=64)], wf,
                       tuning="fastest")


def test_batch_autotune_key_and_persisted_cache(tmp_path, monkeypatch):
    """The batch tuner's cache: keyed by geometry/wall mix, measurement block and
    n_sub class; defaults without an entry; written to and read back from the
    per-architecture file ($DMIPY_AUTOTUNE_CACHE overrides its location)."""
    from dmipy_sim_amd import autotune
    k1 = autotune.batch_key([(3, 1), (2, 2), (3, 1)], [10, 17, 30], 64)
    assert k1 == autotune.batch_key([(2, 2), (3, 1)], [17, 17], 50)
    assert k1 != autotune.batch_key([(2, 2), (3, 1)], [17], 16)
    assert k1 != autotune.batch_key([(2, 2)], [17], 64)
    assert k1 != autotune.batch_key([(2, 2), (3, 1)], [500], 64)
    path = tmp_path / "cache" / "tune.json"
    monkeypatch.setenv("DMIPY_AUTOTUNE_CACHE", str(path))
    assert autotune.default_path() == str(path)
    autotune.clear()
    autotune._loaded["done"] = False
    assert autotune.cached_batch_shape(k1) == autotune.BATCH_DEFAULT
    autotune._cache[k1] = {"shape": [3, 256, 2], "ms": {"4,64,1": 2.0, "3,256,2": 1.0}}
    autotune._save_default()
    autotune.clear()
    assert autotune.cached_batch_shape(k1) == autotune.BATCH_DEFAULT
    autotune._loaded["done"] = False          # a new process: the file is read on first use
    assert autotune.cached_batch_shape(k1) == (3, 256, 2)
    autotune.clear()


def test_autotune_candidates_and_cache(tmp_path):
    from dmipy_sim_amd import autotune
    assert autotune.candidate_blocks(3, False) == [4]
    assert autotune.candidate_blocks(16, False) == [4, 16]
    assert autotune.candidate_blocks(90, False) == [4, 16, 32, 64, 96]
    assert autotune.candidate_blocks(128, True) == [4, 16, 32, 64]
    wf = ds.pgse(0.01, 0.04, 0.05, [[1, 0, 0]], 50)
    plan = ds.make_walk_plan(ds.Cylinder(5e-6, (0, 0, 1)), D, wf.dt)
    autotune.clear()
    assert autotune.cached_block(plan, 128, False) == 0
    key = autotune.plan_key(plan, 128, False)
    autotune._cache[key] = {"meas_block": 64, "ms": {64: 1.0, 128: 2.0}}
    assert autotune.cached_block(plan, 128, False) == 64
    assert autotune.cached_block(plan, 128, True) == 0      # compensation is part of the key
    path = tmp_path / "tune.json"
    autotune.save(str(path))
    autotune.clear()
    autotune.load(str(path))
    assert autotune.cached_block(plan, 128, False) == 64
    autotune.clear()


def test_myelinated_cylinder_unsupported_paths_raise():
    """Full position export and the replay engine are not served for the
    myelinated cylinder (core.py:91, 459-463); the batched launch and the
    walk-once producer refuse it too.  Raised before any GPU work."""
    mc = ds.MyelinatedCylinder(3e-6, 5e-6, [0, 0, 1], D_intra=2e-9, D_extra=2e-9)
    wf = ds.Waveform(G=np.zeros((1, 20, 3), np.float32), dt=1e-4, echo_idx=19)
    with pytest.raises(NotImplementedError, match="return_positions='full'"):
        ds.simulate(100, None, wf, mc, return_positions="full")
    with pytest.raises(NotImplementedError, match="dedicated fused step kernel"):
        ds.simulate(100, None, wf, mc, engine="replay")
    with pytest.raises(ValueError):
        ds.MyelinatedCylinder(5e-6, 3e-6, [0, 0, 1], D_intra=2e-9, D_extra=2e-9)
    assert mc.water_fractions == (1.0, 0.4, 1.0) and mc.D_myelin == 0.0
    assert abs(sum(mc.volume_fraction(c) for c in ("intra", "myelin", "extra")) - 1.0) < 1e-12


def _pmyel_geom(**kw):
    rng = np.random.default_rng(0)
    ri, gr, c = ds.pack_myelinated_cylinders(rng.uniform(1e-6, 2e-6, 8), 0.7, 0.5, seed=1)
    L = float(np.sqrt(np.pi * np.sum((ri / gr) ** 2) / 0.5))
    return ds.PackedMyelinatedCylinders(ri, gr, c, L, N_max=16, **kw)


@pytest.mark.parametrize("kw", [
    dict(),
    dict(D_myelin=3e-10, T2_intra=0.08, T2_extra=0.1, kappa_inner=2e-5, kappa_outer=1e-5),
    dict(rho_outer=2e-5, rho_inner=1e-5, D_extra=[1.5e-9, 2e-9, 2.5e-9, 2e-9, 2e-9, 2e-9, 2e-9, 2e-9],
         T2_myelin=0.015),
], ids=["plain", "perm_t2", "rho_per_axon"])
@pytest.mark.parametrize("dt", [1e-4, 2.5e-5])
def test_packed_myelin_plan_matches_oracle(oracle, kw, dt):
    """Product plan and the oracle's restatement agree on the surface sub-steps, the
    per-axon float32 tables, the per-class dt/T2 and the start boundaries."""
    g = _pmyel_geom(**kw)
    p = ds.make_walk_plan(g, None, dt, T1=0.9)
    o = oracle.oracle_plan(g, None, dt, T1=0.9)
    assert p.n_sub == o["n_sub"] and p.kind == o["kind"] == 11
    for k in ("gamma_dt_sub", "dt_sub", "inv_T1", "nudge", "cell"):
        assert np.float32(getattr(p, k)) == np.float32(o[k]), k
    pm = p.packed_myelin
    np.testing.assert_array_equal(pm["step_len"], o["pm_step"])
    np.testing.assert_array_equal(pm["two_kappa_over_D"], o["pm_k2"])
    np.testing.assert_array_equal(pm["inner"], o["pm_inner"])
    assert [np.float32(v) for v in pm["dt_over_T2"]] == [np.float32(v) for v in o["pm_q2"]]
    assert np.float32(pm["rho_over_D"]) == np.float32(o["pm_rhoD"])
    assert (p.n_sub > 1) == ("rho_outer" in kw)
    for N in (1, 500, 12345):
        n_in, n_my, _ = pm["start_counts"](N)
        mine = np.concatenate([[0], np.cumsum(np.concatenate([n_in, n_my]))])
        np.testing.assert_array_equal(mine, oracle.packed_myelin_start(o, N))


def test_packed_myelin_unsupported_paths_raise():
    g = _pmyel_geom()
    wf = ds.Waveform(G=np.zeros((1, 20, 3), np.float32), dt=1e-4, echo_idx=19)
    with pytest.raises(NotImplementedError, match="r0"):
        ds.simulate(10, None, wf, g, r0=np.zeros((10, 3)))
    with pytest.raises(NotImplementedError, match="return_positions='full'"):
        ds.simulate(10, None, wf, g, return_positions="full")
    with pytest.raises(NotImplementedError, match="PackedMyelinatedCylinders"):
        ds.simulate(10, None, wf, g, engine="replay")
    with pytest.raises(ValueError):
        ds.PackedMyelinatedCylinders([1e-6, 2e-6], 0.7, np.zeros((3, 2)), 1e-5)
    assert abs(sum(g.volume_fraction(c) for c in ("intra", "myelin", "extra")) - 1.0) < 1e-12


def test_float_envelope_rejects_sub_angstrom_lengths():
    """The wall routines' exact unscaled arithmetic holds for lengths >= 1e-10 and
    steps >= 1e-11 in the caller's unit; smaller ones are refused with ValueError
    before any GPU work (FreeDiffusion and Box1D use no square roots)."""
    from dmipy_sim_amd import _native as nat
    wf = ds.Waveform(G=np.zeros((1, 5, 3), np.float32), dt=1e-4, echo_idx=4)
    plan = ds.make_walk_plan(ds.Cylinder(5e-11, (0, 0, 1)), D, wf.dt)
    gd = plan.geometry_desc(n_walkers=64)
    G = np.zeros((1, 5, 3), np.float32)
    wd = nat.Waveform_t(n_meas=1, n_t=5, G=nat.fptr(G), chi=None, n_echo=0, echo_idx=None)
    rd = nat.Run_t(n_walkers=64, seed=0, n_sub=plan.n_sub, step_len=plan.step_len,
                   gamma_dt_sub=plan.gamma_dt_sub, dt_sub=plan.dt_sub, inv_T2=0.0, inv_T1=0.0,
                   has_t2=0, has_t1=0, kahan=0, meas_block=0, r0=None, inj_normals=None,
                   inj_uniforms=None)
    import ctypes as C
    h = C.c_void_p()
    rc = nat.lib().dmw_plan_create(C.byref(h), gd, wd, rd)
    assert rc == 1 and b"float32 range" in nat.lib().dmw_last_error()


def test_recip_rn_matches_ieee_division():
    """The host's correctly rounded reciprocal (the known divisor's RN(1/y) for the
    walk's div_rn_r) equals IEEE 1.0f / y: every mantissa of [1, 2) and 2e6 random
    normal floats (host code of tests/gpu_selftest/arith_check.hip, no GPU call)."""
    import ctypes as C
    lib_path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "gpu_selftest", "libarith_check.so")
    if not os.path.exists(lib_path):
        pytest.fail("tests/gpu_selftest/libarith_check.so is not built (run build())")
    lib = C.CDLL(lib_path)
    lib.arith_check_recip.argtypes = [C.c_uint64, C.c_uint32, C.c_int, C.c_int, C.POINTER(C.c_ulonglong),
                                      C.POINTER(C.c_uint32)]
    bad, first = C.c_ulonglong(0), C.c_uint32(0)
    assert lib.arith_check_recip(2_000_000, 3, 120, 1, C.byref(bad), C.byref(first)) == 0
    assert bad.value == 0, f"{bad.value} mismatches, first y bits {first.value:#x}"


def test_bounce_shortcut_never_skips_a_hit():
    """The walk's bounce-loop shortcut after a reflection (walk_kernel.h no_second_hit:
    inside, heading inward, remaining path <= fl(2|dp| (1 - 2^-20))) only skips
    iterations that find no hit: on 4e6 random float32 post-reflection states (radii
    over two decades, positions nudged inside the wall or anywhere inside, remaining
    paths from 1e-3 R to 3 R, some negative) the iteration's own float32 arithmetic
    (d = -dp + sqrt(max(dp^2 - c, 0)), hit iff EPS < d < rem) never hits where the
    shortcut is taken."""
    rng = np.random.default_rng(17)
    n = 4_000_000
    f = np.float32
    R = (10.0 ** rng.uniform(-6.5, -4.5, n)).astype(f)
    RR = R * R
    ang = rng.uniform(0, 2 * np.pi, n)
    near = rng.random(n) < 0.5
    rad = np.where(near, R * (1.0 - 1e-4 * (1 + rng.random(n))), R * np.sqrt(rng.random(n)))
    x = (rad * np.cos(ang)).astype(f)
    y = (rad * np.sin(ang)).astype(f)
    th = rng.uniform(0, 2 * np.pi, n)
    dx, dy = np.cos(th).astype(f), np.sin(th).astype(f)
    rem = (R * 10.0 ** rng.uniform(-3.0, 0.5, n)).astype(f)
    rem = np.where(rng.random(n) < 0.1, -rem, rem).astype(f)
    dp = dx * x + dy * y
    c = (x * x + y * y) - RR
    take = (dp <= f(-(2.0 ** -60))) & (c < 0) & (rem <= (f(-2.0) * dp) * f(1 - 2.0 ** -20))
    assert take.sum() > n // 4
    disc = np.maximum(dp * dp - c, f(0))
    d = -dp + np.sqrt(disc)
    hit = (d > f(1e-7) * R) & (d < rem)
    assert not np.any(take & hit)


def test_sqrt_cut_is_the_least_square_reaching_r():
    """The walk decides |f| >= R on the squared norm against sqrt_cut(R), the least float
    whose square root reaches R (host code of tests/gpu_selftest/arith_check.hip)."""
    import ctypes as C
    lib_path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "gpu_selftest", "libarith_check.so")
    if not os.path.exists(lib_path):
        pytest.fail("tests/gpu_selftest/libarith_check.so is not built (run build())")
    lib = C.CDLL(lib_path)
    lib.arith_check_sqrt_cut.argtypes = [C.c_uint64, C.c_uint32, C.c_int, C.POINTER(C.c_ulonglong),
                                         C.POINTER(C.c_uint32)]
    bad, first = C.c_ulonglong(0), C.c_uint32(0)
    assert lib.arith_check_sqrt_cut(2_000_000, 5, 40, C.byref(bad), C.byref(first)) == 0
    assert bad.value == 0, f"{bad.value} failures, first r bits {first.value:#x}"


def _valid_plan_args():
    """A valid Cylinder walk description for dmw_plan_create (5 samples, 64 walkers)."""
    from dmipy_sim_amd import _native as nat
    wf = ds.Waveform(G=np.zeros((1, 5, 3), np.float32), dt=1e-4, echo_idx=4)
    plan = ds.make_walk_plan(ds.Cylinder(5e-6, (0, 0, 1)), D, wf.dt)
    gd = plan.geometry_desc(n_walkers=64)
    G = np.zeros((1, 5, 3), np.float32)
    wd = nat.Waveform_t(n_meas=1, n_t=5, G=nat.fptr(G), chi=None, n_echo=0, echo_idx=None)
    rd = nat.Run_t(n_walkers=64, seed=0, n_sub=plan.n_sub, step_len=plan.step_len,
                   gamma_dt_sub=plan.gamma_dt_sub, dt_sub=plan.dt_sub, inv_T2=0.0, inv_T1=0.0,
                   has_t2=0, has_t1=0, kahan=0, meas_block=0, r0=None, inj_normals=None,
                   inj_uniforms=None)
    return gd, wd, rd, G


_EINVAL, _EUNSUP = _native.DMW_EINVAL, _native.DMW_EUNSUPPORTED
_BAD_ECHO_SAME = np.array([3, 3], np.int32)
_BAD_ECHO_RANGE = np.array([5], np.int32)
_INJ = np.zeros(64 * 5 * 3 * 8, np.float32)


@pytest.mark.parametrize("field,value,rc,text", [
    ("run.n_walkers", 0, _EINVAL, "n_walkers"),
    ("run.n_walkers", (1 << 31) + 1, _EINVAL, "n_walkers"),
    ("wave.n_meas", 0, _EINVAL, "empty waveform"),
    ("wave.G", None, _EINVAL, "empty waveform"),
    ("run.n_sub", 0, _EINVAL, "n_sub"),
    ("run.n_sub", 1 << 30, _EINVAL, "too large"),
    ("geom.kind", 12, _EINVAL, "unknown geometry kind"),
    ("geom.wall", 3, _EINVAL, "unknown wall mode"),
    ("geom.kind+wall", (0, 1), _EUNSUP, "FreeDiffusion has no wall"),
    ("geom.kind+wall", (1, 2), _EUNSUP, "Box1D has no permeable wall"),
    ("geom.kind+wall", (8, 1), _EUNSUP, "PermeableShell relaxes only"),
    ("geom.kind+wall", (9, 2), _EINVAL, "0 < r_inner < r_outer"),
    ("geom.kind+wall", (10, 0), _EINVAL, "compartment constants"),
    ("geom.kind+wall", (4, 0), _EINVAL, "packed substrate needs objects"),
    ("wave.echo", _BAD_ECHO_SAME, _EINVAL, "ascending"),
    ("wave.echo", _BAD_ECHO_RANGE, _EINVAL, "ascending"),
    ("wave.n_echo", 1, _EINVAL, "bad echo list"),
    ("inject_without_uniforms", None, _EINVAL, "injected uniforms"),
])
def test_plan_create_argument_checks_before_gpu(field, value, rc, text):
    """dmw_plan_create rejects each malformed description with the status the
    Python layer maps to ValueError / NotImplementedError (core.py:316-345) and a
    message naming the problem, before it touches a GPU."""
    import ctypes as C
    from dmipy_sim_amd import _native as nat
    gd, wd, rd, G = _valid_plan_args()
    if field == "run.n_walkers":
        rd.n_walkers = value
    elif field == "run.n_sub":
        rd.n_sub = value
    elif field == "wave.n_meas":
        wd.n_meas = value
    elif field == "wave.G":
        wd.G = None
    elif field == "geom.kind":
        gd.kind = value
    elif field == "geom.wall":
        gd.wall = value
    elif field == "geom.kind+wall":
        gd.kind, gd.wall = value
        gd.n_obj = 0
    elif field == "wave.echo":
        wd.n_echo = len(value)
        wd.echo_idx = nat.iptr(value)
    elif field == "wave.n_echo":
        wd.n_echo = value
        wd.echo_idx = None
    elif field == "inject_without_uniforms":
        gd.wall = 2
        rd.inj_normals = nat.fptr(_INJ)
    h = C.c_void_p()
    got = nat.lib().dmw_plan_create(C.byref(h), gd, wd, rd)
    assert got == rc, nat.lib().dmw_last_error()
    assert text in nat.lib().dmw_last_error().decode()
    assert not h.value


def test_plan_create_null_arguments():
    import ctypes as C
    from dmipy_sim_amd import _native as nat
    gd, wd, rd, G = _valid_plan_args()
    h = C.c_void_p()
    assert nat.lib().dmw_plan_create(None, gd, wd, rd) == _EINVAL
    assert nat.lib().dmw_plan_create(C.byref(h), None, wd, rd) == _EINVAL
    assert nat.lib().dmw_plan_create(C.byref(h), gd, None, rd) == _EINVAL
    assert nat.lib().dmw_plan_create(C.byref(h), gd, wd, None) == _EINVAL
    assert b"null argument" in nat.lib().dmw_last_error()
    assert nat.lib().dmw_walk(gd, wd, None, None) == _EINVAL


def test_integration_stub_structs_match_header():
    """The ctypes structures of the reference-side binding shown in INTEGRATION.md
    (Option B) have the header's sizes and offsets: a stub missing an appended
    field would let the library write past the caller's structure."""
    import ctypes as C
    txt = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    block = txt.split("```python\n# dmipy_sim/_hip.py", 1)[1].split("\n", 1)[1].split("```", 1)[0]
    defs = block.split("_lib.dmw_walk.argtypes")[0]
    defs = "\n".join(l for l in defs.splitlines() if not l.startswith("_lib = "))
    ns = {}
    exec(defs, ns)
    pairs = {"Geometry": _native.Geometry_t, "Waveform": _native.Waveform_t,
             "Run": _native.Run_t, "Result": _native.Result_t}
    for stub, mirror in pairs.items():
        st = ns[stub]
        assert C.sizeof(st) == C.sizeof(mirror), stub
        assert [f for f, _ in st._fields_] == [f for f, _ in mirror._fields_], stub
        for f, _ in st._fields_:
            assert getattr(st, f).offset == getattr(mirror, f).offset, (stub, f)


def test_unwrap_periodic_restores_a_folded_walk():
    """trajectories.py:82-132: a random walk folded into [-L/2, L/2) on the
    periodic axes is restored (up to the first sample's fold and float32
    rounding); the non-periodic axis is untouched; a save step near L/2 warns."""
    rng = np.random.default_rng(3)
    L = 20e-6
    steps = rng.normal(0, 0.4e-6, (5, 400, 3))
    x = np.cumsum(steps, axis=1) + rng.uniform(-30e-6, 30e-6, (5, 1, 3))
    folded = x.copy()
    folded[..., :2] -= L * np.floor(folded[..., :2] / L + 0.5)
    out = ds.unwrap_periodic(folded.astype(np.float32), L, periodic_axes=(0, 1))
    assert out.dtype == np.float32
    shift = x[:, :1, :2] - folded[:, :1, :2]
    np.testing.assert_allclose(out[..., :2] + shift, x[..., :2], atol=1e-10)
    np.testing.assert_array_equal(out[..., 2], folded[..., 2].astype(np.float32))
    coarse = np.zeros((1, 3, 3), np.float32)
    coarse[0, 1, 0] = 0.47 * L
    with pytest.warns(RuntimeWarning, match="L/2"):
        ds.unwrap_periodic(coarse, L)


def test_gpu_memory_helpers_without_objects():
    """gpu.py:93-185 surface: free_gpu_memory collects and returns the number of
    objects it closed, gpu_session frees on exit, list_gpu_processes returns a
    list (empty without a device or rocm-smi)."""
    assert ds.free_gpu_memory() == 0
    with ds.gpu_session(aggressive=False) as s:
        assert isinstance(s, ds.gpu_session)
    assert isinstance(ds.list_gpu_processes(), list)


def test_batch_over_64_measurements_needs_permeable_packings():
    """More than 64 measurements per waveform: only permeable packings (the
    screened kernel, phase parked in device memory) take them; any other instance
    is refused before anything reaches the device, and so is more than 128."""
    import dmipy_sim_amd as ds
    from substrates import packed_cyl, small_pgse
    wf = small_pgse(n_dirs=23)            # 92 measurements
    with pytest.raises(ValueError, match="permeable packing"):
        ds.BatchedWalk([dict(geometry=packed_cyl(kappa=5e-5), diffusivity=D, n_walkers=64),
                        dict(geometry=ds.Sphere(3e-6), diffusivity=D, n_walkers=64)], wf)
    with pytest.raises(ValueError, match="permeable packing"):
        ds.BatchedWalk([dict(geometry=packed_cyl(), diffusivity=D, n_walkers=64)], wf)
    with pytest.raises(ValueError, match="at most 128"):
        ds.BatchedWalk([dict(geometry=packed_cyl(kappa=5e-5), diffusivity=D, n_walkers=64)],
                       small_pgse(n_dirs=33))
