"""Gradient waveforms: the acquisition side of the walk.

``Waveform.G`` has shape ``(n_measurements, n_t, 3)``, float32, in T/m, sampled on a
uniform grid of step ``dt``.  The constructors keep the reference's integer timing
rules and float32/float64 dtype behaviour so that a waveform built here carries the
same samples as one built by dmipy_sim/waveforms.py:

* ``pgse``      -- dmipy_sim/waveforms.py:93-163
* ``pgste``     -- dmipy_sim/waveforms.py:166-262 (stimulated echo: chi_perp mask, T1 while stored)
* ``ogse``      -- dmipy_sim/waveforms.py:265-350
* ``trapezoidal_ogse`` / ``b_trapezoidal_ogse`` -- dmipy_sim/waveforms.py:353-487
* ``cpmg``      -- dmipy_sim/waveforms.py:851-917
* ``calc_b``    -- dmipy_sim/waveforms.py:490-516
* ``calc_btensor`` / ``btensor_invariants`` -- dmipy_sim/waveforms.py:519-595
* ``ste`` / ``pte`` -- dmipy_sim/waveforms.py:598-720 (spherical / planar tensor encoding)
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


def pgste(delta, TM, G_magnitude, bvecs, n_t, slew_rate=DEFAULT_SLEW_RATE):
    """Pulsed-gradient stimulated echo (dmipy_sim/waveforms.py:166-262).

    Total duration ``2*delta + TM`` over ``n_t`` samples: a dephasing lobe on
    samples ``[0, n_pulse)``, the mixing time (magnetisation stored along z: no
    gradient, ``chi_perp`` 0, only T1 acts) until ``n_pulse + round(TM/dt)``, then
    the negated rephasing lobe.  ``n_pulse = max(1, round(delta/dt))``; the ramp
    length follows the peak amplitude passed here.  The stimulated echo keeps half
    the magnetisation (``stimulated_echo=True``: ``simulate`` scales by 0.5).
    """
    bvecs, g, n_meas = _as_meas(bvecs, G_magnitude)
    dt = (2.0 * delta + TM) / (n_t - 1)
    n_pulse = max(1, round(delta / dt))
    i_recall = n_pulse + round(TM / dt)
    _, square = resolve_slew(slew_rate)
    n_rise = _lobe_n_rise(square, float(np.max(np.abs(g))), slew_rate, dt, n_pulse)
    n_second = min(n_pulse, n_t - i_recall)
    G = np.zeros((n_meas, n_t, 3), dtype=np.float32)
    G_disp = np.zeros((n_meas, n_t, 3), dtype=np.float32)
    for m in range(n_meas):
        amp = g[m] * bvecs[m]
        _fill_lobe(G, m, 0, n_pulse, amp, n_rise)
        _fill_lobe(G, m, i_recall, n_second, -amp, n_rise)
        _fill_lobe(G_disp, m, 0, n_pulse, amp, n_rise)
        _fill_lobe(G_disp, m, i_recall, n_second, amp, n_rise)
    chi_perp = np.ones(n_t, dtype=bool)
    chi_perp[n_pulse:i_recall] = False
    rf_events = [{'t_s': 0.0, 'label': 'Mz→Mxy', 'flip_deg': 90},
                 {'t_s': n_pulse * dt, 'label': 'store', 'flip_deg': 90},
                 {'t_s': i_recall * dt, 'label': 'recall', 'flip_deg': 90}]
    return Waveform(G=G, dt=float(dt), echo_idx=n_t - 1, rf_events=rf_events,
                    G_display=G_disp, chi_perp=chi_perp, TM=float(TM), stimulated_echo=True)


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


def trapezoidal_ogse(N, delta, DELTA, G_magnitude, bvecs, n_t, rise_time=None, slew_rate=200e3):
    """Trapezoidal N-lobe OGSE (dmipy_sim/waveforms.py:353-455; Drobnjak et al.,
    MRM 2016).  Two blocks of N lobes of ``n_lobe = max(1, round(delta/(N dt)))``
    samples each, the first at sample 0 with signs +, -, +, ..., the second at
    ``round(DELTA/dt)`` with the signs flipped (spin-echo convention); every lobe
    ramps linearly over ``min(round(rise_time/dt), n_lobe//2)`` samples, with
    ``rise_time = max |G| / slew_rate`` when not given.  N = 1 is a trapezoidal PGSE.
    """
    bvecs, g, n_meas = _as_meas(bvecs, G_magnitude)
    g_peak = float(np.max(np.abs(g)))
    if rise_time is None:
        rise_time = g_peak / slew_rate if g_peak > 0 else 0.0
    dt = (DELTA + delta) / (n_t - 1)
    n_lobe = max(1, round(delta / (N * dt)))
    n_rise = min(max(0, round(rise_time / dt)), n_lobe // 2)
    n_DELTA = round(DELTA / dt)
    G = np.zeros((n_meas, n_t, 3), dtype=np.float32)
    for m in range(n_meas):
        for block0, first_sign in ((0, 1.0), (n_DELTA, -1.0)):
            for k in range(N):
                amp = np.float32(first_sign * (-1.0) ** k) * g[m]
                i0 = block0 + k * n_lobe
                i1 = i0 + n_lobe
                if n_rise > 0:
                    up = np.linspace(0.0, amp, n_rise, endpoint=False)
                    dn = np.linspace(amp, 0.0, n_rise, endpoint=False)
                    G[m, i0:i0 + n_rise, :] = up[:, None] * bvecs[m]
                    G[m, i0 + n_rise:i1 - n_rise, :] = amp * bvecs[m]
                    G[m, i1 - n_rise:i1, :] = dn[:, None] * bvecs[m]
                else:
                    G[m, i0:i1, :] = amp * bvecs[m]
    gap_mid = (n_lobe * N + n_DELTA) // 2
    rf_events = [{'t_s': 0.0, 'label': 'Mz→Mxy', 'flip_deg': 90},
                 {'t_s': gap_mid * dt, 'label': 'refocus', 'flip_deg': 180}]
    return Waveform(G=G, dt=float(dt), echo_idx=n_t - 1, rf_events=rf_events)


def b_trapezoidal_ogse(N, delta, DELTA, G, rise_time):
    """Closed-form b-value of the trapezoidal N-lobe OGSE in s/m^2 (Eq. 2 of
    Drobnjak et al., MRM 2016; dmipy_sim/waveforms.py:458-487): with the ramp
    fraction x = rise_time N / delta, the within-block term
    2 G^2 gamma^2 delta^3 / (15 N^2) (5 - 15x/2 - 5x^2/4 + 4x^3), plus for odd N the
    net moment of a block, (delta - N rise_time) / N, acting over DELTA - delta."""
    x = rise_time * N / delta
    within = (2.0 * G ** 2 * GAMMA ** 2 * delta ** 3 / (15.0 * N ** 2)
              * (5.0 - 7.5 * x - 1.25 * x ** 2 + 4.0 * x ** 3))
    net = (1 - (-1) ** N) * (delta - N * rise_time) / (2.0 * N)
    return float(within + G ** 2 * GAMMA ** 2 * (DELTA - delta) * net ** 2)


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


def calc_btensor(waveform):
    """B-tensor per measurement, B_ij = int q_i(t) q_j(t) dt in s/m^2
    (dmipy_sim/waveforms.py:519-543): the q(t) and integration rules of calc_b,
    so that its trace equals calc_b to float64 rounding.  (n_meas, 3, 3) float64."""
    G = np.asarray(waveform.G, dtype=np.float32)
    dt = waveform.dt
    q = np.cumsum(G * dt, axis=1) * GAMMA
    B = np.trapezoid(q[:, :, :, None] * q[:, :, None, :], dx=dt, axis=1)
    return np.asarray(B).astype(np.float64)


def btensor_invariants(B):
    """(b, b_delta, b_eta) of B-tensors (n_meas, 3, 3) (dmipy_sim/waveforms.py:546-595).

    b = trace; with eigenvalues l1 >= l2 >= l3, a prolate tensor (l1 farther from
    b/3 than l3) has b_delta = (l1 - (l2 + l3)/2) / b and b_eta = (l2 - l3) / b, an
    oblate one b_delta = (l3 - (l1 + l2)/2) / b and b_eta = (l1 - l2) / b: linear
    encoding 1, spherical 0, planar -1/2 (Szczepankiewicz et al. 2021).  b = 0 rows
    give zero shape parameters."""
    lam = np.linalg.eigvalsh(np.asarray(B))[:, ::-1]
    l1, l2, l3 = lam[:, 0], lam[:, 1], lam[:, 2]
    b = l1 + l2 + l3
    bs = np.where(b > 0, b, 1.0)
    prolate = np.abs(l1 - b / 3.0) >= np.abs(l3 - b / 3.0)
    b_delta = np.where(prolate, l1 - 0.5 * (l2 + l3), l3 - 0.5 * (l1 + l2)) / bs
    b_eta = np.where(prolate, l2 - l3, l1 - l2) / bs
    return b, b_delta, b_eta


def _bipolar_axes(axes, T_total, G_magnitude, n_t, slew_rate):
    """One bipolar pair per axis, back to back: the total time is cut into
    2 len(axes) segments of round(n_t / (2 len(axes))) samples, axis i encodes
    (+) on segment 2i and decodes (-) on segment 2i+1, so every axis's q returns
    to zero before the next starts (dmipy_sim/waveforms.py:598-720)."""
    dt = T_total / (n_t - 1)
    n_seg = max(1, round(n_t / (2 * len(axes))))
    _, square = resolve_slew(slew_rate)
    n_rise = _lobe_n_rise(square, abs(float(G_magnitude)), slew_rate, dt, n_seg)
    G = np.zeros((1, n_t, 3), dtype=np.float32)
    for i, axis in enumerate(axes):
        start = 2 * i * n_seg
        mid = min((2 * i + 1) * n_seg, n_t)
        end = min((2 * i + 2) * n_seg, n_t)
        amp = (G_magnitude * np.asarray(axis)).astype(np.float32)
        _fill_lobe(G, 0, start, mid - start, amp, n_rise)
        _fill_lobe(G, 0, mid, end - mid, -amp, n_rise)
    return Waveform(G=G, dt=float(dt), echo_idx=n_t - 1,
                    rf_events=[{'t_s': 0.0, 'label': 'Mz→Mxy', 'flip_deg': 90}])


def ste(delta, DELTA, G_magnitude, n_t, slew_rate=DEFAULT_SLEW_RATE):
    """Spherical tensor encoding, B = (b/3) I (dmipy_sim/waveforms.py:598-655):
    bipolar pairs along x, y and z in turn over DELTA + delta (only the sum is
    used).  One measurement; scale with set_b."""
    return _bipolar_axes(np.eye(3, dtype=np.float32), DELTA + delta, G_magnitude, n_t, slew_rate)


def pte(delta, DELTA, G_magnitude, plane_normal, n_t, slew_rate=DEFAULT_SLEW_RATE):
    """Planar tensor encoding, eigenvalues (b/2, b/2, 0) with the zero along
    ``plane_normal`` (dmipy_sim/waveforms.py:658-720): bipolar pairs along two
    orthonormal in-plane axes u, v in turn (u from Gram-Schmidt of x, or of y when
    the normal is close to x; v = n x u)."""
    n = np.asarray(plane_normal, dtype=np.float64)
    n = n / np.linalg.norm(n)
    ref = np.array([1.0, 0.0, 0.0]) if abs(n[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
    u = ref - np.dot(ref, n) * n
    u /= np.linalg.norm(u)
    v = np.cross(n, u)
    v /= np.linalg.norm(v)
    return _bipolar_axes([u, v], DELTA + delta, G_magnitude, n_t, slew_rate)


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
