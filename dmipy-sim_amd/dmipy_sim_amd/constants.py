"""Physical constants shared by the waveform builders and the walk.

Mirrors dmipy_sim/constants.py:6-30 (GAMMA, DEFAULT_SLEW_RATE, resolve_slew).
"""
import numpy as np

#: Proton gyromagnetic ratio in rad/(s*T) (dmipy_sim/constants.py:6).
GAMMA = 267.513e6

#: Default gradient slew rate in T/m/s for the slew-limited constructors
#: (dmipy_sim/constants.py:14).
DEFAULT_SLEW_RATE = 200.0


def resolve_slew(slew_rate):
    """Validate ``slew_rate``; return ``(value, is_instantaneous)``.

    A positive rate in T/m/s gives slew-limited (trapezoidal) lobes, ``np.inf``
    the square limit; ``None`` or a non-positive value raises ``ValueError``
    (dmipy_sim/constants.py:17-30).
    """
    if slew_rate is None or not (slew_rate > 0):
        raise ValueError(
            "slew_rate must be a positive number (T/m/s) for a realistic "
            "slew-limited gradient, or np.inf for the idealized instantaneous "
            "(square) limit; got {!r}.".format(slew_rate))
    return slew_rate, (not np.isfinite(slew_rate))
