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

For batched launches three knobs are tuned, one after the other (coordinate
descent from the defaults): the persistent wavefronts per SIMD (bounded by the
kernel's register occupancy), the walkers a wavefront takes per queue pop (64,
128, 256 or 512, walked as consecutive 64-walker tiles: fewer pops and
instance loads against a coarser tail) and the unrolling of the sub-step loop
(1 or 2).  Each is timed on the batch itself with every walker count scaled
down, and the choice is kept per (geometry/wall mix, measurement block, n_sub
class) key.  None of the knobs changes a result: every tile walks its own
Philox streams and writes its own partial row.

The cache lives in memory and, by default, in a JSON file per GPU
architecture ($DMIPY_AUTOTUNE_CACHE, else
~/.cache/dmipy_sim_amd/autotune-<arch>.json), read on first use and written
after every tuning run.
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
    _load_default()
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
    _save_default()
    return best, times


BATCH_WAVES_PER_SIMD = (1, 2, 3, 4)
BATCH_WALKERS_PER_WAVE = (64, 128, 256, 512)
BATCH_UNROLL = (1, 2)
BATCH_DEFAULT = (0, 0, 0)       # the library's defaults: 4 waves/SIMD, 64 walkers, no unrolling


def batch_key(kind_walls, n_subs, n_meas):
    """Cache key of a batch: its (geometry kind, wall) mix, measurement block
    (16 or 64) and the n_sub class of its median instance."""
    mix = ",".join(sorted({f"{int(k)}.{int(w)}" for k, w in kind_walls}))
    return f"batch:{mix}:{16 if n_meas <= 16 else 64}:{_nsub_class(int(np.median(n_subs)))}"


def tune_batch(instances, waveforms, walker_scale=0.02, reps=2, log=None):
    """Launch shape of a batch by coordinate descent: persistent waves per SIMD,
    then walkers per wavefront pop, then sub-step unrolling, each candidate
    timed (median of ``reps`` graph launches after one warm launch) on the same
    batch with every walker count scaled by ``walker_scale`` (at least one
    64-walker tile per instance).  Returns ((waves_per_simd, walkers_per_wave,
    unroll), {shape: ms}) and caches the choice under the batch's key."""
    from .batch import BatchedWalk
    _load_default()
    scaled = []
    for s in instances:
        d = dict(s)
        d["n_walkers"] = max(64, int(s["n_walkers"] * walker_scale))
        scaled.append(d)
    times = {}

    def timed(shape):
        if shape not in times:
            bw = BatchedWalk(scaled, waveforms, waves_per_simd=shape[0], walkers_per_wave=shape[1],
                             unroll=shape[2])
            bw.run(seed=1)
            times[shape] = float(np.median([bw.run(seed=2 + r)[1]["total_ms"] for r in range(reps)]))
            key = bw.key
            bw.close()
            timed.key = key
            if log:
                log(f"autotune batch: {shape[0]} waves/SIMD, {shape[1]} walkers/wave, unroll "
                    f"{shape[2]} -> {times[shape]:.2f} ms")
        return times[shape]

    best = (4, 64, 1)
    for axis, cands in ((0, BATCH_WAVES_PER_SIMD), (1, BATCH_WALKERS_PER_WAVE), (2, BATCH_UNROLL)):
        for c in cands:
            shape = best[:axis] + (c,) + best[axis + 1:]
            if timed(shape) < timed(best):
                best = shape
    _cache[timed.key] = {"shape": list(best),
                         "ms": {",".join(map(str, k)): v for k, v in times.items()}}
    _save_default()
    return best, times


def cached_batch_shape(key):
    """(waves_per_simd, walkers_per_wave, unroll) tuned for a batch key, or the
    library's defaults (zeros) if none."""
    _load_default()
    e = _cache.get(key)
    return BATCH_DEFAULT if e is None or "shape" not in e else tuple(int(v) for v in e["shape"])


def cached_block(plan, n_meas, kahan):
    """The tuned block for a plan, or 0 (the built-in heuristic) if none."""
    _load_default()
    e = _cache.get(plan_key(plan, n_meas, kahan))
    return 0 if e is None or "meas_block" not in e else int(e["meas_block"])


def _serialise(e):
    if "shape" in e:
        return {"shape": list(e["shape"]), "ms": dict(e["ms"])}
    return {"meas_block": e["meas_block"], "ms": {str(b): t for b, t in e["ms"].items()}}


def _parse(v):
    if "shape" in v:
        return {"shape": [int(x) for x in v["shape"]], "ms": {k: float(t) for k, t in v["ms"].items()}}
    return {"meas_block": int(v["meas_block"]), "ms": {int(b): float(t) for b, t in v["ms"].items()}}


def save(path):
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump({k: _serialise(v) for k, v in _cache.items()}, f, indent=1)
    os.replace(tmp, path)


def load(path):
    with open(path) as f:
        data = json.load(f)
    for k, v in data.items():
        _cache[k] = _parse(v)


def default_path():
    """$DMIPY_AUTOTUNE_CACHE, else ~/.cache/dmipy_sim_amd/autotune-<arch>.json
    for the current device's architecture (None without a device)."""
    env = os.environ.get("DMIPY_AUTOTUNE_CACHE")
    if env:
        return env
    from . import _native as nat
    arch = nat.device_arch()
    if not arch:
        return None
    return os.path.join(os.path.expanduser("~"), ".cache", "dmipy_sim_amd", f"autotune-{arch}.json")


_loaded = {"path": None, "done": False}


def _load_default():
    if _loaded["done"]:
        return
    _loaded["done"] = True
    try:
        path = default_path()
    except Exception:
        path = None
    _loaded["path"] = path
    if path and os.path.exists(path):
        try:
            load(path)
        except (OSError, ValueError, KeyError):
            pass            # an unreadable cache is re-tuned, not fatal


def _save_default():
    path = _loaded["path"]
    if path:
        try:
            save(path)
        except OSError:
            pass            # a read-only home keeps the in-memory cache only


def clear():
    """Forget every tuned choice (in memory; the file stays until the next save)."""
    _cache.clear()
    _loaded["done"] = True
