"""Device guard (dmipy_sim/gpu.py:50-87 semantics).

The walk only runs on the GPU; there is no CPU path to fall back to.  The guard
keeps the reference's contract for ``require_gpu``: ``True`` raises without a
device, ``None`` warns for large runs (>= LARGE_RUN_WALKERS walkers), ``False``
stays silent.  A run without a device then fails loudly in the walk itself.
"""
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
