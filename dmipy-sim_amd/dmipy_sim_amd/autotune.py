"""Launch-shape autotuner for the walk.

The walk kernel keeps one walker per lane (64 walkers per wavefront) and the
phase of ``meas_block`` measurements per pass in registers or matrix
accumulators; a walk with n_meas measurements runs ceil(n_meas / meas_block)
passes (grid.y), each re-walking the same trajectories.  Narrow blocks leave
registers free for more resident wavefronts; wide blocks walk each trajectory
fewer times.  The best width depends on the geometry's per-step cost, n_sub
and n_meas, so it is measured: every valid width is timed on a probe (a
walker subset over the first samples of the waveform) and the fastest is kept
per (geometry kind, wall mode, n_meas, n_sub class, compensation) key.

For batched launches the tuned knob is the number of persistent wavefronts per
SIMD (bounded by the kernel's register occupancy), timed on the batch with
every walker count scaled down.
"""
import json
import os

import numpy as np

from .physics import make_walk_plan

MEAS_BLOCKS = (4, 16, 32, 64, 96, 128)
_cache = {}


def _nsub_class(n_sub):
    return int(np.ceil(np.log2(max(1, n_sub))))


def plan_key(plan, n_meas, kahan):
    return f"walk:{plan.kind}:{plan.wall}:{n_meas}:{_nsub_class(plan.n_sub)}:{int(bool(kahan))}"


def candidate_blocks(n_meas, kahan):
    """Widths worth timing: every block up to the first one covering n_meas
    (96 and 128 need the uncompensated matrix-core form)."""
    out = []
    for mb in MEAS_BLOCKS:
        if kahan and mb > 64:
            break
        out.append(mb)
        if mb >= n_meas:
            break
    return out


def tune_walk(geometry, diffusivity, waveform, n_walkers, T2=None, T1=None, kahan=None,
              probe_walkers=131072, probe_samples=32, reps=3, log=None):
    """Time every candidate measurement block on a probe and return the fastest
    (cached).  Returns (meas_block, {block: kernel ms})."""
    from .core import PlannedWalk, _kahan_auto
    from .waveforms import Waveform
    G = np.asarray(waveform.G, np.float32)
    n_meas, n_t = G.shape[:2]
    if kahan is None:
        kahan = _kahan_auto(n_t, n_meas)
    plan = make_walk_plan(geometry, diffusivity, waveform.dt, T2=T2, T1=T1)
    key = plan_key(plan, n_meas, kahan)
    if key in _cache:
        return _cache[key]["meas_block"], _cache[key]["ms"]
    ts = min(n_t, probe_samples)
    probe = Waveform(G=G[:, :ts].copy(), dt=waveform.dt, echo_idx=ts - 1)
    chi = getattr(waveform, "chi_perp", None)
    if chi is not None:
        probe.chi_perp = np.asarray(chi, np.float32)[:ts]
    n = int(min(n_walkers, probe_walkers))
    times = {}
    for mb in candidate_blocks(n_meas, kahan):
        pw = PlannedWalk(geometry, diffusivity, probe, n, T2=T2, T1=T1, kahan=kahan, meas_block=mb)
        pw.run(seed=1)
        times[mb] = float(np.median([pw.run(seed=2 + r)["kernel_ms"] for r in range(reps)]))
        pw.close()
        if log:
            log(f"autotune {key}: meas_block {mb} -> {times[mb]:.3f} ms")
    best = min(times, key=times.get)
    _cache[key] = {"meas_block": best, "ms": times}
    return best, times


def tune_batch(instances, waveforms, walker_scale=0.02, candidates=(1, 2, 3, 4), log=None):
    """Persistent wavefronts per SIMD for a batch: each candidate timed on the
    same batch with every walker count scaled by ``walker_scale`` (at least one
    64-walker unit per instance).  Returns (waves_per_simd, {candidate: ms})."""
    from .batch import BatchedWalk
    scaled = []
    for s in instances:
        d = dict(s)
        d["n_walkers"] = max(64, int(s["n_walkers"] * walker_scale))
        scaled.append(d)
    times = {}
    for wps in candidates:
        bw = BatchedWalk(scaled, waveforms, waves_per_simd=wps)
        bw.run(seed=1)
        times[wps] = float(np.median([bw.run(seed=2 + r)[1]["total_ms"] for r in range(2)]))
        bw.close()
        if log:
            log(f"autotune batch: {wps} waves/SIMD -> {times[wps]:.2f} ms")
    best = min(times, key=times.get)
    return best, times


def cached_block(plan, n_meas, kahan):
    """The tuned block for a plan, or 0 (the built-in heuristic) if none."""
    e = _cache.get(plan_key(plan, n_meas, kahan))
    return 0 if e is None else int(e["meas_block"])


def save(path):
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(path, "w") as f:
        json.dump({k: {"meas_block": v["meas_block"], "ms": {str(b): t for b, t in v["ms"].items()}}
                   for k, v in _cache.items()}, f, indent=1)


def load(path):
    with open(path) as f:
        data = json.load(f)
    for k, v in data.items():
        _cache[k] = {"meas_block": int(v["meas_block"]),
                     "ms": {int(b): float(t) for b, t in v["ms"].items()}}


def clear():
    _cache.clear()
