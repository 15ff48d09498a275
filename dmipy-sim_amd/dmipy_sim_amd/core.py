"""Entry points of the walk: ``simulate``, ``simulate_cpmg``, ``simulate_mixture``.

Signatures, argument checks, walker chunking, OOM back-off, output shapes and
dtypes follow dmipy_sim/core.py:193-969.  The walk itself -- walker keys,
start positions, the n_t x n_sub step scan, the wall physics, relaxation and the
ensemble reduction (core.py:436-735) -- is one call of the HIP library through
``walk()``.

Engine routing (core.py:22-190): ``engine='auto'`` and ``'fused'`` run the
fused-engine semantics (the reference's oracle engine); ``'replay'`` checks the
same gap list as the reference and then walks with the walk-once producer's
sub-step rule (R/6), accumulating the phase on the fly instead of storing the
trajectory.
"""
import warnings

import numpy as np

from . import _native as nat
from . import geometries as geo
from .gpu import check_gpu
from . import autotune
from .physics import make_walk_plan, WALL_PERMEATE

_REPLAY_AUTO_GEOM_NAMES = frozenset({
    "FreeDiffusion", "Box1D", "Sphere", "Cylinder", "PackedCylinders", "PackedSpheres",
})


def _replay_gap(geometry, *, return_positions, return_compartments,
                return_walker_signals, diffusivity):
    """Why the replay engine cannot serve this run, or None (core.py:77-106)."""
    if return_positions is not False:
        return "return_positions is a fused single-pass internal (no replay equivalent)"
    if return_compartments is not False:
        return "return_compartments is a fused single-pass internal (no replay equivalent)"
    if return_walker_signals:
        return "return_walker_signals is a fused single-pass internal (no replay equivalent)"
    if getattr(geometry, '_is_myelinated', False):
        return "MyelinatedCylinder uses a dedicated fused step kernel (no replay walk)"
    if getattr(geometry, '_is_packed_myelinated', False):
        return ("PackedMyelinatedCylinders fused single-reflection kernel is not "
                "position-parity with the multi-bounce replay walk")
    if getattr(geometry, 'permeability', None) is not None:
        return ("membrane permeability is single-pass walk semantics "
                "(fused-only; scalar replay has no exchange reservoir)")
    if diffusivity is None:
        return "replay needs an explicit diffusivity for the sub-step auto-tune"
    return None


def _kahan_auto(n_t, n_meas):
    """Compensated (Kahan) phase for long waveforms (>= 1500 samples) whose
    measurements fit one 64-wide pass; otherwise float32 accumulation, whose
    rounding error (<= n_t ulp(|phi|), ~1e-4 rad RMS at 2000 samples and
    |phi| ~ 100 rad) changes signals by far less than 1e-5."""
    return n_t >= 1500 and n_meas <= 64


def walk(geometry, diffusivity, waveform, n_walkers, seed=123, T2=None, T1=None, r0=None,
         *, outputs=(), inject_normals=None, inject_uniforms=None, kahan=None,
         meas_block=0, substep_rule="fused", echo_indices=None, counters=False, plan=None,
         first_walker=0, n_total=None, noalloc=False, slices=None):
    """One walk on the GPU; returns a dict of numpy arrays.

    ``outputs`` selects optional per-walker arrays: 'final_pos', 'comp_origin',
    'comp_final', 'comp_full', 'pos_full', 'n_cross', 'log_w', 'phi', 'signal_im'.
    Positions are returned in the lab frame.  ``inject_normals`` (N, n_t*n_sub, 3)
    and ``inject_uniforms`` (N, n_t*n_sub) replace the Philox stream with given
    step increments (raw normal triples, normalised in the kernel as the reference
    does, physics.py:299-301).  ``first_walker`` makes this call walk walkers
    [first_walker, first_walker + n_walkers) of a larger walk (their random streams
    and start positions): a shard of a walk split over calls or devices;
    ``n_total`` is then the whole walk's walker count (default first_walker +
    n_walkers), which fixes the myelinated substrates' start allotment.
    ``noalloc``: run through dmw_walk_noalloc (no device allocation; the call's
    device buffers must fit the library's 64 MiB scratch array).
    ``slices``: sample counts of successive dmw_plan_advance calls that walk the
    waveform in pieces, the walkers kept in device memory between them (the last
    piece is extended to n_t); the result equals the one-call walk's.
    """
    G = np.asarray(waveform.G, dtype=np.float32)
    if G.ndim != 3 or G.shape[2] != 3:
        raise ValueError("waveform.G must have shape (n_measurements, n_t, 3)")
    n_meas, n_t, _ = G.shape
    if plan is None:
        plan = make_walk_plan(geometry, diffusivity, waveform.dt, T2=T2, T1=T1,
                              substep_rule=substep_rule)
    N = int(n_walkers)
    if N < 1:
        raise ValueError("n_walkers must be >= 1")
    rot = plan.rot
    if rot is not None:
        G = np.ascontiguousarray(np.einsum('ij,mtj->mti', rot.astype(np.float64), G),
                                 dtype=np.float32)
    chi = getattr(waveform, 'chi_perp', None)
    chi = None if chi is None else np.ascontiguousarray(np.asarray(chi, dtype=np.float32).reshape(n_t))
    echo = None if echo_indices is None else np.ascontiguousarray(echo_indices, dtype=np.int32)

    if r0 is not None:
        r0 = np.asarray(r0, dtype=np.float32).reshape(N, 3)
        if rot is not None:
            r0 = (rot.astype(np.float64) @ r0.T.astype(np.float64)).T
        r0 = np.ascontiguousarray(r0, dtype=np.float32)
    S = n_t * plan.n_sub
    if inject_normals is not None:
        inject_normals = np.ascontiguousarray(inject_normals, dtype=np.float32)
        if inject_normals.shape != (N, S, 3):
            raise ValueError(f"inject_normals must have shape {(N, S, 3)}")
        if rot is not None:
            inject_normals = np.ascontiguousarray(
                np.einsum('ij,wsj->wsi', rot.astype(np.float64), inject_normals), dtype=np.float32)
        if plan.wall == WALL_PERMEATE or plan.myelin is not None or plan.packed_myelin is not None:
            if inject_uniforms is None:
                raise ValueError("inject_uniforms required on a permeable wall")
            inject_uniforms = np.ascontiguousarray(inject_uniforms, dtype=np.float32).reshape(N, S)
        else:
            inject_uniforms = None
    if kahan is None:
        kahan = _kahan_auto(n_t, n_meas)
    if not meas_block:
        meas_block = autotune.cached_block(plan, n_meas, kahan)

    if int(first_walker) < 0:
        raise ValueError("first_walker must be >= 0")
    total = int(first_walker) + N if n_total is None else int(n_total)
    if total < int(first_walker) + N:
        raise ValueError("n_total must cover walkers [first_walker, first_walker + n_walkers)")
    gd = plan.geometry_desc(n_walkers=total)
    G = np.ascontiguousarray(G, dtype=np.float32)   # referenced by wd until the call returns
    wd = nat.Waveform_t(n_meas=n_meas, n_t=n_t, G=nat.fptr(G),
                        chi=nat.fptr(chi), n_echo=0 if echo is None else len(echo),
                        echo_idx=nat.iptr(echo))
    rd = nat.Run_t(n_walkers=N, seed=int(seed) & (2**64 - 1), n_sub=plan.n_sub,
                   step_len=plan.step_len, gamma_dt_sub=plan.gamma_dt_sub, dt_sub=plan.dt_sub,
                   inv_T2=plan.inv_T2, inv_T1=plan.inv_T1, has_t2=int(plan.has_t2),
                   has_t1=int(plan.has_t1), kahan=int(bool(kahan)), meas_block=int(meas_block),
                   r0=nat.fptr(r0), inj_normals=nat.fptr(inject_normals),
                   inj_uniforms=nat.fptr(inject_uniforms), first_walker=int(first_walker))
    rows = 1 if echo is None else len(echo)
    res = {"signal": np.zeros(rows * n_meas, np.float64)}
    want = set(outputs)
    if "signal_im" in want:
        res["signal_im"] = np.zeros(rows * n_meas, np.float64)
    shapes = {"final_pos": ((N, 3), np.float32), "comp_origin": ((N,), np.int32),
              "comp_final": ((N,), np.int32), "comp_full": ((N, n_t), np.int32),
              "pos_full": ((N, n_t, 3), np.float32), "n_cross": ((N,), np.uint32),
              "log_w": ((N,), np.float32), "phi": ((N, n_meas), np.float32)}
    for k, (shp, dt) in shapes.items():
        if k in want:
            res[k] = np.zeros(shp, dt)
    out = nat.Result_t(signal=nat.dptr(res["signal"]), signal_im=nat.dptr(res.get("signal_im")),
                       final_pos=nat.fptr(res.get("final_pos")),
                       comp_origin=nat.iptr(res.get("comp_origin")),
                       comp_final=nat.iptr(res.get("comp_final")),
                       comp_full=nat.iptr(res.get("comp_full")),
                       pos_full=nat.fptr(res.get("pos_full")),
                       n_cross=nat.uptr(res.get("n_cross")), log_w=nat.fptr(res.get("log_w")),
                       phi=nat.fptr(res.get("phi")), want_counters=int(bool(counters)))
    if slices is not None:
        if noalloc:
            raise ValueError("slices and noalloc exclude each other")
        _walk_in_slices(gd, wd, rd, out, [int(k) for k in slices], n_t)
    else:
        entry = nat.lib().dmw_walk_noalloc if noalloc else nat.lib().dmw_walk
        nat.check(entry(gd, wd, rd, out))
    if echo is not None:
        res["signal"] = res["signal"].reshape(rows, n_meas)
        if "signal_im" in res:
            res["signal_im"] = res["signal_im"].reshape(rows, n_meas)
    if rot is not None:
        rinv = rot.T.astype(np.float64)
        for k in ("final_pos", "pos_full"):
            if k in res:
                res[k] = np.einsum('ij,...j->...i', rinv, res[k].astype(np.float64)).astype(np.float32)
    res["kernel_ms"] = out.kernel_ms
    res["phase_form"] = nat.PHASE_FORMS[out.phase_form]
    res["n_sub"] = plan.n_sub
    res["plan"] = plan
    if counters:
        res["n_hits"] = int(out.n_hits)
        res["n_cand"] = int(out.n_cand)
    return res


def _walk_in_slices(gd, wd, rd, out, slices, n_t):
    """One walk through dmw_plan_advance, ``slices`` samples per call; the last
    call walks every remaining sample.  out.kernel_ms is summed over the calls."""
    import ctypes as C
    lib = nat.lib()
    h = C.c_void_p()
    nat.check(lib.dmw_plan_create(C.byref(h), gd, wd, rd))
    try:
        t_done = C.c_int32(0)
        total_ms = 0.0
        for k in slices:
            if k < 1:
                raise ValueError("slices must be positive sample counts")
            nat.check(lib.dmw_plan_advance(h, rd.seed, k, out, C.byref(t_done)))
            total_ms += out.kernel_ms
            if t_done.value == n_t:
                break
        if t_done.value < n_t:
            nat.check(lib.dmw_plan_advance(h, rd.seed, n_t - t_done.value, out, C.byref(t_done)))
            total_ms += out.kernel_ms
        out.kernel_ms = total_ms
    finally:
        lib.dmw_plan_destroy(h)


def simulate(
    n_walkers: int,
    diffusivity=None,
    waveform=None,
    geometry=None,
    seed: int = 123,
    T2: float = None,
    T1: float = None,
    return_positions=False,
    return_compartments=False,
    return_walker_signals: bool = False,
    r0=None,
    walker_batch_size: int = None,
    require_gpu=None,
    engine: str = "auto",
    _allow_oom_backoff: bool = True,
):
    """Monte-Carlo diffusion signal (dmipy_sim/core.py:193-762).

    Returns float32 signals of shape (n_measurements,), plus, on request, final or
    per-sample positions, (origin, final|full) compartments and per-walker
    (log_w, phi), in the reference's order.
    """
    if return_compartments not in (False, 'final', 'full'):
        raise ValueError("return_compartments must be False, 'final', or 'full'; "
                         f"got {return_compartments!r}")
    if return_positions not in (False, True, 'full'):
        raise ValueError("return_positions must be False, True, or 'full'; "
                         f"got {return_positions!r}")
    if engine not in ("auto", "replay", "fused"):
        raise ValueError(f"engine must be 'auto', 'replay', or 'fused'; got {engine!r}")
    substep_rule = "fused"
    if engine == "replay":
        gap = _replay_gap(geometry, return_positions=return_positions,
                          return_compartments=return_compartments,
                          return_walker_signals=return_walker_signals,
                          diffusivity=diffusivity)
        if gap is not None:
            raise NotImplementedError(f"engine='replay' cannot serve this run: {gap}. "
                                      "Use engine='fused' (or 'auto').")
        # Routed before the fused engine's OOM back-off and walker batching, as the
        # reference does (core.py:334-349): one producer walk under one seed.  The
        # trajectories stay resident in HBM, so walker_batch_size (the reference's
        # host-side producer batch) bounds nothing here and is accepted as is.
        check_gpu(n_walkers, require_gpu, what="simulate")
        wf_r = waveform.waveform if hasattr(waveform, 'waveform') else waveform
        return _simulate_via_replay(n_walkers, diffusivity, wf_r, geometry, seed=seed,
                                    T2=T2, T1=T1, r0=r0)
    _myelinated = (getattr(geometry, '_is_myelinated', False)
                   or getattr(geometry, '_is_packed_myelinated', False))
    if return_positions == 'full' and _myelinated:
        raise NotImplementedError(   # core.py:459-463
            "return_positions='full' is supported for standard geometries "
            "(including Mesh), not MyelinatedCylinder / PackedMyelinatedCylinders.")
    if r0 is not None and getattr(geometry, '_is_packed_myelinated', False):
        raise NotImplementedError(   # core.py:542-546
            "simulate(r0=...) is unsupported for PackedMyelinatedCylinders; it "
            "initialises walkers (and their compartments) from `seed` via init_positions.")

    check_gpu(n_walkers, require_gpu, what="simulate")

    if _allow_oom_backoff:
        bs = walker_batch_size
        while True:
            try:
                return simulate(n_walkers, diffusivity=diffusivity, waveform=waveform,
                                geometry=geometry, seed=seed, T2=T2, T1=T1,
                                return_positions=return_positions,
                                return_compartments=return_compartments,
                                return_walker_signals=return_walker_signals, r0=r0,
                                walker_batch_size=bs, require_gpu=False, engine=engine,
                                _allow_oom_backoff=False)
            except RuntimeError as exc:
                m = str(exc)
                if not ('RESOURCE_EXHAUSTED' in m or 'out of memory' in m.lower()):
                    raise
                cur = bs if bs is not None else n_walkers
                nxt = cur // 2
                if nxt < 1:
                    raise
                warnings.warn("simulate() hit GPU OOM at walker_batch_size={}; retrying at "
                              "{}.".format(cur, nxt), RuntimeWarning, stacklevel=2)
                bs = nxt

    if walker_batch_size is not None and walker_batch_size < n_walkers:
        return _simulate_in_walker_batches(
            n_walkers, walker_batch_size, seed=seed, diffusivity=diffusivity,
            waveform=waveform, geometry=geometry, T2=T2, T1=T1, r0=r0,
            return_positions=return_positions, return_compartments=return_compartments,
            return_walker_signals=return_walker_signals, engine=engine)

    if hasattr(waveform, 'waveform'):
        waveform = waveform.waveform
    track = return_compartments is not False
    outs = []
    if return_positions == 'full':
        outs.append('pos_full')
    elif return_positions:
        outs.append('final_pos')
    if track:
        outs.append('comp_origin')
        outs.append('comp_full' if return_compartments == 'full' else 'comp_final')
    if return_walker_signals:
        outs += ['log_w', 'phi']
    res = walk(geometry, diffusivity, waveform, n_walkers, seed=seed, T2=T2, T1=T1, r0=r0,
               outputs=outs, substep_rule=substep_rule)
    sig = res["signal"].astype(np.float32)
    if getattr(waveform, 'stimulated_echo', False):
        sig = sig * np.float32(0.5)
    result = [sig]
    if return_positions == 'full':
        result.append(res['pos_full'])
    elif return_positions:
        result.append(res['final_pos'])
    if track:
        result.append(res['comp_origin'])
        result.append(res['comp_full'] if return_compartments == 'full' else res['comp_final'])
    if return_walker_signals:
        result.append(res['log_w'])
        result.append(res['phi'])
    return result[0] if len(result) == 1 else tuple(result)


def _simulate_via_replay(n_walkers, diffusivity, waveform, geometry, *, seed, T2, T1, r0):
    """Signal via the walk-once path (dmipy_sim/core.py:115-190): the producer
    walk saves float16 positions (and, with surface relaxivity, the boundary
    log-weight at rho/D = 1); the waveform, T2/T1 and rho are then applied by
    the float64 replay contraction.  Trajectories never leave the GPU."""
    from .trajectories import TrajectorySet
    G = np.asarray(waveform.G, dtype=np.float32)
    dt = float(waveform.dt)
    n_t = G.shape[1]
    chi_perp = getattr(waveform, 'chi_perp', None)
    chi = (np.asarray(chi_perp, dtype=np.float64).reshape(n_t)
           if chi_perp is not None else np.ones(n_t, dtype=np.float64))
    rho = getattr(geometry, 'surface_relaxivity_t2', None)
    has_surf = (rho is not None and float(rho) > 0.0
                and not isinstance(geometry, geo.FreeDiffusion))
    ts = TrajectorySet.simulate(n_walkers, diffusivity, geometry, dt * (n_t - 1), dt, seed=seed,
                                save_relaxation_data=has_surf, r0=r0)
    try:
        sig = ts.replay(G, dt, chi_perp=chi, T2=T2, T1=T1,
                        surface_relaxivity=float(rho) if has_surf else None,
                        D=float(diffusivity),
                        stimulated_echo=bool(getattr(waveform, 'stimulated_echo', False)))
    finally:
        ts.close()
    return np.asarray(sig, dtype=np.float32)


def _simulate_in_walker_batches(n_walkers, walker_batch_size, *, seed, diffusivity, waveform,
                                geometry, T2, T1, r0, return_positions, return_compartments,
                                return_walker_signals, engine="fused"):
    """Chunks of ``walker_batch_size`` walkers with seeds seed+1+b, recombined as a
    size-weighted mean; per-walker outputs concatenate (core.py:765-818)."""
    n_batches = (n_walkers + walker_batch_size - 1) // walker_batch_size
    track = return_compartments is not False
    acc = None
    pos_l, origin_l, comp_l, lw_l, phi_l = [], [], [], [], []
    for b in range(n_batches):
        start = b * walker_batch_size
        end = min(start + walker_batch_size, n_walkers)
        nb = end - start
        print(f"  simulate: walkers {start}–{end - 1} "
              f"({int(100 * end / n_walkers)}%)...", flush=True)
        out = simulate(nb, diffusivity=diffusivity, waveform=waveform, geometry=geometry,
                       seed=seed + 1 + b, T2=T2, T1=T1,
                       r0=(None if r0 is None else np.asarray(r0)[start:end]),
                       return_positions=return_positions,
                       return_compartments=return_compartments,
                       return_walker_signals=return_walker_signals,
                       walker_batch_size=None, require_gpu=False, engine=engine,
                       _allow_oom_backoff=False)
        items = list(out) if isinstance(out, tuple) else [out]
        sig = np.asarray(items.pop(0))
        acc = sig * nb if acc is None else acc + sig * nb
        if return_positions:
            pos_l.append(np.asarray(items.pop(0)))
        if track:
            origin_l.append(np.asarray(items.pop(0)))
            comp_l.append(np.asarray(items.pop(0)))
        if return_walker_signals:
            lw_l.append(np.asarray(items.pop(0)))
            phi_l.append(np.asarray(items.pop(0)))
    result = [acc / n_walkers]
    if return_positions:
        result.append(np.concatenate(pos_l, axis=0))
    if track:
        result.append(np.concatenate(origin_l, axis=0))
        result.append(np.concatenate(comp_l, axis=0))
    if return_walker_signals:
        result.append(np.concatenate(lw_l, axis=0))
        result.append(np.concatenate(phi_l, axis=0))
    return result[0] if len(result) == 1 else tuple(result)


def simulate_mixture(compartments, waveform, seed=123):
    """Volume-fraction-weighted sum of independent compartments, compartment i
    walked with seed + i (core.py:821-858).

    All compartments run in one batched persistent launch (each as instance
    word 0 with seed offset i, i.e. the walk ``simulate(..., seed=seed + i)``
    performs) when the waveform has at most 64 measurements and needs no
    compensated phase and no compartment is myelinated (those walk on the
    single-walk path only); otherwise one ``simulate`` call per compartment."""
    total = sum(c['fraction'] for c in compartments)
    if abs(total - 1.0) > 1e-6:
        raise ValueError(f"Compartment fractions must sum to 1, got {total}")
    n_meas, n_t = np.asarray(waveform.G).shape[:2]
    batchable = all(getattr(c['geometry'], '_walk_kind', None)
                    not in (geo.KIND_MYELIN, geo.KIND_PACKED_MYELIN) for c in compartments)
    if compartments and batchable and n_meas <= 64 and not _kahan_auto(n_t, n_meas):
        from .batch import BatchedWalk
        check_gpu(None)
        specs = [dict(geometry=c['geometry'], diffusivity=c['diffusivity'],
                      n_walkers=int(c['n_walkers']), stream=0, seed_offset=i)
                 for i, c in enumerate(compartments)]
        bw = BatchedWalk(specs, waveform)
        try:
            sig, _ = bw.run(seed)
        finally:
            bw.close()
        signal = None
        for i, comp in enumerate(compartments):
            w = comp['fraction'] * sig[i].astype(np.float32)
            signal = w if signal is None else signal + w
        return signal
    signal = None
    for i, comp in enumerate(compartments):
        s = simulate(n_walkers=comp['n_walkers'], diffusivity=comp['diffusivity'],
                     waveform=waveform, geometry=comp['geometry'], seed=seed + i)
        w = comp['fraction'] * s
        signal = w if signal is None else signal + w
    return signal


def simulate_cpmg(n_walkers, diffusivity, waveform, geometry, *, T2=None, seed=123,
                  walker_batch_size=None, require_gpu=None):
    """Multi-echo CPMG from one walk: the ensemble signal sampled at every
    ``waveform.echo_indices`` (core.py:861-969).  Returns float32
    (n_echoes, n_measurements)."""
    check_gpu(n_walkers, require_gpu, what="simulate_cpmg")
    if hasattr(waveform, 'waveform'):
        waveform = waveform.waveform
    echo_indices = getattr(waveform, 'echo_indices', None)
    if echo_indices is None:
        raise ValueError("simulate_cpmg needs a multi-echo waveform with echo_indices set; "
                         "build it with dmipy_sim_amd.cpmg(...).")
    echo_indices = np.asarray(echo_indices, dtype=int)
    if walker_batch_size is not None and walker_batch_size < n_walkers:
        n_batches = (n_walkers + walker_batch_size - 1) // walker_batch_size
        acc = None
        for b in range(n_batches):
            start = b * walker_batch_size
            nb = min(walker_batch_size, n_walkers - start)
            print(f"  simulate_cpmg: walkers {start}–{start + nb - 1} "
                  f"({int(100 * (start + nb) / n_walkers)}%)...", flush=True)
            s = simulate_cpmg(nb, diffusivity, waveform, geometry, T2=T2, seed=seed + 1 + b,
                              walker_batch_size=None, require_gpu=False)
            acc = s * nb if acc is None else acc + s * nb
        return acc / n_walkers
    # The reference's echo slice of the per-step trace; repeated indices repeat rows.
    uniq, inv = np.unique(echo_indices, return_inverse=True)
    res = walk(geometry, diffusivity, waveform, n_walkers, seed=seed, T2=T2,
               echo_indices=uniq)
    return res["signal"][inv].astype(np.float32)


class PlannedWalk:
    """A walk whose inputs (waveform, geometry tables) stay resident on the GPU;
    ``run(seed)`` launches the walk and the reduction only.  Used for throughput
    measurement and for repeated walks of one configuration with new seeds.
    ``meas_block``: measurements per kernel pass (0: the tuned value if one is
    cached, else the built-in choice; "tune": time the candidates now)."""

    def __init__(self, geometry, diffusivity, waveform, n_walkers, T2=None, T1=None,
                 kahan=None, meas_block=0, substep_rule="fused"):
        import ctypes as C
        G = np.asarray(waveform.G, dtype=np.float32)
        n_meas, n_t, _ = G.shape
        self.plan = make_walk_plan(geometry, diffusivity, waveform.dt, T2=T2, T1=T1,
                                   substep_rule=substep_rule)
        plan = self.plan
        if plan.rot is not None:
            G = np.einsum('ij,mtj->mti', plan.rot.astype(np.float64), G)
        self._G = np.ascontiguousarray(G, dtype=np.float32)
        chi = getattr(waveform, 'chi_perp', None)
        self._chi = None if chi is None else np.ascontiguousarray(chi, dtype=np.float32).reshape(n_t)
        if kahan is None:
            kahan = _kahan_auto(n_t, n_meas)
        if meas_block == "tune":
            meas_block, _ = autotune.tune_walk(geometry, diffusivity, waveform, n_walkers, T2=T2,
                                               T1=T1, kahan=kahan)
        elif not meas_block:
            meas_block = autotune.cached_block(plan, n_meas, kahan)
        self.meas_block = int(meas_block)
        gd = plan.geometry_desc(n_walkers=int(n_walkers))
        wd = nat.Waveform_t(n_meas=n_meas, n_t=n_t, G=nat.fptr(self._G), chi=nat.fptr(self._chi),
                            n_echo=0, echo_idx=None)
        rd = nat.Run_t(n_walkers=int(n_walkers), seed=0, n_sub=plan.n_sub,
                       step_len=plan.step_len, gamma_dt_sub=plan.gamma_dt_sub,
                       dt_sub=plan.dt_sub, inv_T2=plan.inv_T2, inv_T1=plan.inv_T1,
                       has_t2=int(plan.has_t2), has_t1=int(plan.has_t1),
                       kahan=int(bool(kahan)), meas_block=int(meas_block),
                       r0=None, inj_normals=None, inj_uniforms=None)
        h = C.c_void_p()
        nat.check(nat.lib().dmw_plan_create(C.byref(h), gd, wd, rd))
        self._h = h
        nat.track(self)
        self.n_meas, self.n_t, self.n_walkers = n_meas, n_t, int(n_walkers)
        self.walker_steps = int(n_walkers) * n_t * plan.n_sub
        self.kahan = bool(kahan)

    def run(self, seed=123, counters=False):
        sig = np.zeros(self.n_meas, np.float64)
        out = nat.Result_t(signal=nat.dptr(sig), want_counters=int(bool(counters)))
        nat.check(nat.lib().dmw_plan_execute(self._h, int(seed) & (2**64 - 1), out))
        res = {"signal": sig, "kernel_ms": out.kernel_ms, "phase_form": nat.PHASE_FORMS[out.phase_form]}
        if counters:
            res["n_hits"] = int(out.n_hits)
            res["n_cand"] = int(out.n_cand)
        return res

    def advance(self, n_samples, seed=123, counters=False):
        """Walk the next ``n_samples`` samples of the walk in progress (a new walk
        with ``seed`` when none is); the walkers stay in device memory between
        calls.  Returns ``t_done`` and ``kernel_ms`` of this call, and once
        ``t_done == n_t`` the same entries as ``run`` (counters summed over the
        calls).  Needs the parked phase form unless one call walks every sample."""
        import ctypes as C
        sig = np.zeros(self.n_meas, np.float64)
        out = nat.Result_t(signal=nat.dptr(sig), want_counters=int(bool(counters)))
        t_done = C.c_int32(0)
        nat.check(nat.lib().dmw_plan_advance(self._h, int(seed) & (2**64 - 1), int(n_samples), out,
                                              C.byref(t_done)))
        res = {"t_done": int(t_done.value), "kernel_ms": out.kernel_ms,
               "phase_form": nat.PHASE_FORMS[out.phase_form]}
        if res["t_done"] == self.n_t:
            res["signal"] = sig
            if counters:
                res["n_hits"] = int(out.n_hits)
                res["n_cand"] = int(out.n_cand)
        return res

    def rewind(self):
        """Abandon the walk in progress."""
        nat.check(nat.lib().dmw_plan_rewind(self._h))

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            nat.lib().dmw_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
