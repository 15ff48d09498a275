"""Gradient waveforms: the acquisition side of the walk.

``Waveform.G`` has shape ``(n_measurements, n_t, 3)``, float32, in T/m, sampled on a
uniform grid of step ``dt``.  The constructors keep the reference's integer timing
rules and float32/float64 dtype behaviour so that a waveform built here carries the
same samples as one built by dmipy_sim/waveforms.py:

* ``pgse``      -- dmipy_sim/waveforms.py:93-163
* ``ogse``      -- dmipy_sim/waveforms.py:265-350
* ``cpmg``      -- dmipy_sim/waveforms.py:851-917
* ``calc_b``    -- dmipy_sim/waveforms.py:490-516
* ``set_b``     -- dmipy_sim/waveforms.py:723-772
* ``rotate_waveform`` / ``tile_waveform`` -- dmipy_sim/waveforms.py:775-848

They are host-side numpy; the walk consumes ``G`` once per call.
"""
from dataclasses import dataclass
from warnings import warn

import numpy as np

from .constants import GAMMA, DEFAULT_SLEW_RATE, resolve_slew


@dataclass
class Waveform:
    """Gradient waveform for the Monte-Carlo walk (dmipy_sim/waveforms.py:39-90).

    Attributes
    ----------
    G : np.ndarray, (n_measurements, n_t, 3) float32, T/m.
    dt : float, uniform time step in seconds.
    echo_idx : int, sample index of the echo (last sample for a spin echo).
    rf_events : list, visual markers only.
    G_display : np.ndarray, physical (same-sign) gradient for display only.
    echo_indices : np.ndarray, echo sample indices of a multi-echo train (CPMG).
    chi_perp : np.ndarray, (n_t,) transverse-coherence flag (1 transverse, 0 stored).
    TM : float, mixing time of a stimulated echo.
    stimulated_echo : bool, apply the 0.5 stimulated-echo amplitude factor.
    """
    G: np.ndarray
    dt: float
    echo_idx: int
    rf_events: list = None
    G_display: np.ndarray = None
    echo_indices: np.ndarray = None
    chi_perp: np.ndarray = None
    TM: float = None
    stimulated_echo: bool = False


def _fill_lobe(arr, m, i0, n_pulse, amp_vec, n_rise):
    """Write one lobe of peak ``amp_vec`` into samples ``[i0, i0+n_pulse)`` of
    measurement ``m``: a rectangle when ``n_rise == 0``, otherwise linear ramps of
    ``n_rise`` samples around a flat top (dmipy_sim/waveforms.py:15-29)."""
    end = i0 + n_pulse
    n_rise = min(int(n_rise), n_pulse // 2)
    if n_rise <= 0:
        arr[m, i0:end, :] = amp_vec
        return
    ramp_up = np.linspace(0.0, 1.0, n_rise, endpoint=False)
    ramp_dn = np.linspace(1.0, 0.0, n_rise, endpoint=False)
    arr[m, i0:i0 + n_rise, :] = ramp_up[:, None] * amp_vec
    arr[m, i0 + n_rise:end - n_rise, :] = amp_vec
    arr[m, end - n_rise:end, :] = ramp_dn[:, None] * amp_vec


def _lobe_n_rise(square, G_peak, slew_rate, dt, n_pulse):
    """Ramp length in samples of a slew-limited lobe, 0 in the square limit
    (dmipy_sim/waveforms.py:32-36)."""
    if square or G_peak <= 0:
        return 0
    return min(max(0, round((G_peak / float(slew_rate)) / dt)), n_pulse // 2)


def _as_meas(bvecs, G_magnitude):
    bvecs = np.asarray(bvecs, dtype=np.float32)
    if bvecs.ndim == 1:
        bvecs = bvecs[None, :]
    n_meas = bvecs.shape[0]
    g = np.broadcast_to(np.asarray(G_magnitude, dtype=np.float32), (n_meas,))
    return bvecs, g, n_meas


def pgse(delta, DELTA, G_magnitude, bvecs, n_t, slew_rate=DEFAULT_SLEW_RATE):
    """Pulsed-gradient spin echo (dmipy_sim/waveforms.py:93-163).

    Total duration ``DELTA + delta`` over ``n_t`` samples, ``dt = T/(n_t-1)``;
    lobes start at samples 0 and ``round(DELTA/dt)`` and last
    ``max(1, round(delta/dt))`` samples; the second lobe is negated (bipolar
    convention encoding the 180 refocusing).  The ramp length is set by the peak
    amplitude passed here, before any ``set_b`` rescaling.
    """
    bvecs, g, n_meas = _as_meas(bvecs, G_magnitude)
    dt = (DELTA + delta) / (n_t - 1)
    n_pulse = max(1, round(delta / dt))
    n_DELTA = round(DELTA / dt)
    _, square = resolve_slew(slew_rate)
    n_rise = _lobe_n_rise(square, float(np.max(np.abs(g))), slew_rate, dt, n_pulse)

    G = np.zeros((n_meas, n_t, 3), dtype=np.float32)
    G_disp = np.zeros((n_meas, n_t, 3), dtype=np.float32)
    for m in range(n_meas):
        amp = g[m] * bvecs[m]
        _fill_lobe(G, m, 0, n_pulse, amp, n_rise)
        _fill_lobe(G, m, n_DELTA, n_pulse, -amp, n_rise)
        _fill_lobe(G_disp, m, 0, n_pulse, amp, n_rise)
        _fill_lobe(G_disp, m, n_DELTA, n_pulse, amp, n_rise)

    refocus = (n_pulse + n_DELTA) // 2
    rf_events = [{'t_s': 0.0, 'label': 'Mz→Mxy', 'flip_deg': 90},
                 {'t_s': refocus * dt, 'label': 'refocus', 'flip_deg': 180}]
    return Waveform(G=G, dt=float(dt), echo_idx=n_t - 1, rf_events=rf_events,
                    G_display=G_disp)


def ogse(frequency, T_total, G_magnitude, bvecs, n_t, kind='cosine'):
    """Oscillating-gradient spin echo (dmipy_sim/waveforms.py:265-350).

    The frequency is snapped to an integer number ``N`` of periods per half echo
    (warning when the snap exceeds 10 %); the second half is the negated time
    mirror of the first, and the middle sample is zeroed for odd ``n_t``, so the
    gradient moment nulls exactly at the echo.
    """
    bvecs, g, n_meas = _as_meas(bvecs, G_magnitude)
    dt = T_total / (n_t - 1)
    t = np.arange(n_t, dtype=np.float64) * dt
    half = T_total / 2.0
    n_per = max(1, int(round(frequency * half)))
    f_eff = n_per / half
    if abs(f_eff - frequency) > 0.1 * max(frequency, 1e-9):
        warn(f"ogse: frequency snapped {frequency:.1f} -> {f_eff:.1f} Hz "
             f"({n_per} periods per half-echo of {half*1e3:.1f} ms) for a valid OGSE")
    if kind == 'cosine':
        env = np.cos(2 * np.pi * f_eff * t)
    elif kind == 'sine':
        env = np.sin(2 * np.pi * f_eff * t)
    else:
        raise ValueError(f"ogse kind must be 'cosine' or 'sine', got {kind!r}")
    mid = n_t // 2
    env[n_t - mid:] = -env[:mid][::-1]
    if n_t % 2 == 1:
        env[mid] = 0.0

    G = np.zeros((n_meas, n_t, 3), dtype=np.float32)
    for m in range(n_meas):
        G[m] = (g[m] * env)[:, None] * bvecs[m]
    refocus = (n_t - 1) // 2
    rf_events = [{'t_s': 0.0, 'label': 'Mz→Mxy', 'flip_deg': 90},
                 {'t_s': refocus * dt, 'label': 'refocus', 'flip_deg': 180}]
    return Waveform(G=G, dt=float(dt), echo_idx=n_t - 1, rf_events=rf_events)


def cpmg(n_echoes, TE, G_magnitude, bvecs, n_t_per_echo=100):
    """Instantaneous-pulse CPMG echo train (dmipy_sim/waveforms.py:851-917).

    ``n_t = n_echoes*n_t_per_echo + 1`` samples over ``n_echoes*TE``; the effective
    gradient flips sign at every refocusing time ``(k+1/2)*TE``; echoes sit at
    sample ``k*n_t_per_echo``, k = 1..n_echoes (``echo_indices``).
    """
    bvecs, g, n_meas = _as_meas(bvecs, G_magnitude)
    n_t = n_echoes * n_t_per_echo + 1
    dt = (n_echoes * TE) / (n_t - 1)
    t = np.arange(n_t) * dt
    sign = np.ones(n_t, dtype=np.float32)
    for k in range(n_echoes):
        sign[t >= (k + 0.5) * TE] *= -1.0
    G_phys = np.zeros((n_meas, n_t, 3), dtype=np.float32)
    G_eff = np.zeros((n_meas, n_t, 3), dtype=np.float32)
    for m in range(n_meas):
        vec = g[m] * bvecs[m]
        G_phys[m] = vec
        G_eff[m] = sign[:, None] * vec
    rf_events = [{'t_s': 0.0, 'label': 'Mz→Mxy', 'flip_deg': 90}]
    rf_events += [{'t_s': (k + 0.5) * TE, 'label': 'refocus', 'flip_deg': 180}
                  for k in range(n_echoes)]
    echo_indices = np.arange(1, n_echoes + 1) * n_t_per_echo
    return Waveform(G=G_eff, dt=float(dt), echo_idx=n_t - 1, rf_events=rf_events,
                    G_display=G_phys, echo_indices=echo_indices)


def calc_b(waveform):
    """b-value per measurement in s/m^2 (dmipy_sim/waveforms.py:490-516).

    q(t) = GAMMA * cumsum(G*dt) (left-point rule, float32), b = trapezoid(|q|^2, dt),
    returned as float64.  The left-point q matches the walk's right-endpoint phase
    rule, so ``set_b`` targets the b-value the walk actually applies.
    """
    G = np.asarray(waveform.G, dtype=np.float32)
    dt = waveform.dt
    q = np.cumsum(G * dt, axis=1) * GAMMA
    q2 = np.sum(q ** 2, axis=2)
    b = np.trapezoid(q2, dx=dt, axis=1)
    return np.asarray(b).astype(np.float64)


def set_b(waveform, b_target):
    """Scale each measurement to the target b-value in s/m^2
    (dmipy_sim/waveforms.py:723-772).  Warns when the targets look like s/mm^2."""
    b_now = calc_b(waveform)
    b_arr = np.asarray(b_target, dtype=np.float64).ravel()
    nonzero = b_arr[b_arr > 0]
    if nonzero.size > 0 and nonzero.max() < 1e5:
        warn(f"set_b: b_target values appear very small (max={nonzero.max():.4g} s/m²). "
             "Did you pass b-values in s/mm² instead of s/m²? "
             "Multiply by 1e6 to convert: set_b(wf, b_mm2 * 1e6). "
             "Typical SI b-values are 1e8–3e9 s/m² (100–3000 s/mm²).",
             UserWarning, stacklevel=2)
    b_t = np.broadcast_to(np.asarray(b_target, dtype=np.float64), b_now.shape)
    scale = np.sqrt(b_t / b_now).astype(np.float32)
    G_new = (np.asarray(waveform.G, dtype=np.float32) * scale[:, None, None]).astype(np.float32)
    G_disp = waveform.G_display
    if G_disp is not None:
        G_disp = (np.asarray(G_disp) * scale[:, None, None]).astype(np.float32)
    return Waveform(G=G_new, dt=waveform.dt, echo_idx=waveform.echo_idx,
                    rf_events=waveform.rf_events, G_display=G_disp,
                    echo_indices=waveform.echo_indices,
                    chi_perp=waveform.chi_perp, TM=waveform.TM,
                    stimulated_echo=waveform.stimulated_echo)


def rotate_waveform(waveform, R=None, *, theta=None):
    """Rotate the gradient vectors, ``G_new = G @ R.T``; ``theta`` rotates about y
    taking z to (sin theta, 0, cos theta) (dmipy_sim/waveforms.py:775-818)."""
    if (R is None) == (theta is None):
        raise ValueError("rotate_waveform: pass exactly one of R or theta.")
    if theta is not None:
        c, s = np.cos(float(theta)), np.sin(float(theta))
        R = np.array([[c, 0.0, s], [0.0, 1.0, 0.0], [-s, 0.0, c]], dtype=np.float32)
    R = np.asarray(R, dtype=np.float32)
    G = np.asarray(waveform.G, dtype=np.float32)
    G_rot = np.einsum('mtj,ij->mti', G, R).astype(np.float32)
    return Waveform(G=G_rot, dt=waveform.dt, echo_idx=waveform.echo_idx,
                    rf_events=waveform.rf_events, echo_indices=waveform.echo_indices,
                    chi_perp=waveform.chi_perp,
                    TM=waveform.TM, stimulated_echo=waveform.stimulated_echo)


def tile_waveform(waveform, n_copies):
    """Stack ``n_copies`` copies along the measurement axis
    (dmipy_sim/waveforms.py:821-848)."""
    G = np.tile(np.asarray(waveform.G, dtype=np.float32), (n_copies, 1, 1))
    return Waveform(G=G, dt=waveform.dt, echo_idx=waveform.echo_idx,
                    rf_events=waveform.rf_events, echo_indices=waveform.echo_indices,
                    chi_perp=waveform.chi_perp,
                    TM=waveform.TM, stimulated_echo=waveform.stimulated_echo)


def fibonacci_directions(n):
    """``n`` near-uniform unit vectors on the sphere (Fibonacci lattice):
    z_i = 1 - (2i+1)/n, azimuth_i = i*pi*(3 - sqrt 5)."""
    i = np.arange(n, dtype=np.float64)
    z = 1.0 - (2.0 * i + 1.0) / n
    az = i * np.pi * (3.0 - np.sqrt(5.0))
    s = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    return np.stack([s * np.cos(az), s * np.sin(az), z], axis=1)
