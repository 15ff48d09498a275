"""Device guard and device-memory helpers (dmipy_sim/gpu.py).

The walk only runs on the GPU; there is no CPU path to fall back to.  The guard
keeps the reference's contract for ``require_gpu``: ``True`` raises without a
device, ``None`` warns for large runs (>= LARGE_RUN_WALKERS walkers), ``False``
stays silent.  A run without a device then fails loudly in the walk itself.

Device memory is owned by the library's objects (planned walks, batched walks,
surrogate networks, trajectory sets): each frees its memory on ``close()`` or when
collected.  ``free_gpu_memory`` collects unreferenced ones and, when asked, closes
every live one; ``gpu_session`` does that on exit; ``list_gpu_processes`` reports
who holds the GPU.
"""
import gc
import json
import subprocess
import warnings

from . import _native

LARGE_RUN_WALKERS = 10_000


def gpu_available() -> bool:
    """True when the HIP walk library sees at least one device."""
    try:
        return _native.device_count() > 0
    except _native.NativeLibraryMissing:
        return False
    except OSError:
        return False


def check_gpu(n_walkers=None, require_gpu=None, what="MC simulation") -> bool:
    """Guard against running without a GPU; returns whether a GPU is visible."""
    on_gpu = gpu_available()
    if require_gpu is False or on_gpu:
        return on_gpu
    msg = (f"{what}: no GPU visible to the HIP walk library; the Monte-Carlo walk runs "
           f"only on the GPU. Pass require_gpu=False to silence this warning.")
    if require_gpu is True:
        raise RuntimeError(msg)
    if n_walkers is None or n_walkers >= LARGE_RUN_WALKERS:
        warnings.warn(msg, RuntimeWarning, stacklevel=3)
    return False


def free_gpu_memory(aggressive: bool = False, clear_compilation_cache: bool = True) -> int:
    """Release device memory held by this process (dmipy_sim/gpu.py:93-134).

    The default is safe while objects are in use: a garbage-collection pass frees
    the device memory of every library object no longer referenced.  With
    ``aggressive`` every live planned walk, batched walk, network and trajectory
    set is closed as well (using one afterwards raises).  There is no compilation
    cache to clear (the kernels are compiled ahead of time), so
    ``clear_compilation_cache`` is accepted for the reference's signature and has
    no effect.  Returns the number of objects closed (0 unless ``aggressive``)."""
    del clear_compilation_cache
    gc.collect()
    closed = 0
    if aggressive:
        for obj in _native.live_objects():
            if getattr(obj, "_h", None) is not None:
                try:
                    obj.close()
                    closed += 1
                except Exception:
                    pass
        gc.collect()
    return closed


class gpu_session:
    """Context manager that frees device memory on exit (dmipy_sim/gpu.py:137-157)."""

    def __init__(self, aggressive: bool = False, clear_compilation_cache: bool = True):
        self.aggressive = aggressive
        self.clear_compilation_cache = clear_compilation_cache

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        free_gpu_memory(aggressive=self.aggressive,
                        clear_compilation_cache=self.clear_compilation_cache)
        return False


def list_gpu_processes():
    """Best-effort list of processes holding GPU memory, from ``rocm-smi
    --showpids --json`` (dmipy_sim/gpu.py:160-185, which asks nvidia-smi): dicts
    ``{pid, used_mib, name}``, empty when the tool or a device is unavailable.
    Reporting only: nothing is ever killed."""
    try:
        out = subprocess.run(["rocm-smi", "--showpids", "--json"], capture_output=True,
                             text=True, timeout=10)
        data = json.loads(out.stdout or "{}")
    except Exception:
        return []
    procs = []
    for key, val in (data.get("system", {}) or {}).items():
        if not str(key).upper().startswith("PID"):
            continue
        pid = str(key)[3:]
        fields = [f.strip() for f in str(val).split(",")]
        if not pid.isdigit():
            continue
        name = fields[0] if fields else ""
        used = None
        if len(fields) >= 3 and fields[2].isdigit():
            used = int(fields[2]) // (1024 * 1024)
        procs.append({"pid": int(pid), "used_mib": used, "name": name})
    return procs
