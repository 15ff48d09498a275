"""CPU restatement of the walk-once / replay path (test infrastructure only).

Follows the reference:
  producer constants   dmipy_sim/core.py:1084-1148 (n_t, dt, sub-step rule R/6 or
                       R/25 with R = radius | sphere_radius | length | min radius,
                       float32 step length, rho/D = 1 when relaxation data are saved)
  replay               dmipy_sim/trajectories.py:501-641 (waveform resampling,
                       nearest-sample chi, float64 phase contraction, T2/T1/rho
                       log-weights, mean of exp(log_w) cos(phi))
in float64 numpy.  Pinned in tests/test_replay_oracle.py by closed forms:
zero-moment waveforms on static walkers, summation-by-parts phases of uniformly
moving walkers, exp(-dt/T2 sum chi - dt/T1 sum (1-chi)), constant boundary
log-weights, linear-waveform resampling, and the sub-step rule's definition.
"""
import math

import numpy as np

GAMMA = 267.513e6
f32 = np.float32


def producer_constants(geometry_radius, permeable, D, T_max, dt_save, save_relax):
    """n_t, dt_actual, sub_steps, dt_sim, float32 step length, rho/D of the walls."""
    n_t = int(round(T_max / dt_save)) + 1
    dt = T_max / (n_t - 1)
    if geometry_radius is None:
        sub = 1
    else:
        div = 3750.0 if permeable else 216.0
        sub = max(1, int(math.ceil(dt / (geometry_radius ** 2 / (div * D)))))
    dt_sim = dt / sub
    step = float(np.sqrt(np.float32(6.0 * D * dt_sim)))
    return dict(n_t=n_t, dt=dt, sub_steps=sub, dt_sim=dt_sim, step_len=step,
                rho_over_D=1.0 if save_relax else 0.0)


def resample(G, dt_wf, dt_traj, n_t_traj, chi):
    """G on the trajectory grid (linear, zero outside) and chi (nearest sample,
    zero past the waveform end); chi is (n_t_wf,) or (n_meas, n_t_wf)."""
    G = np.asarray(G, np.float32)
    n_meas, n_t_wf, _ = G.shape
    t_wf = np.arange(n_t_wf, dtype=np.float64) * dt_wf
    t_tr = np.arange(n_t_traj, dtype=np.float64) * dt_traj
    if abs(dt_wf - dt_traj) / max(abs(dt_traj), 1e-30) > 1e-9:
        Gt = np.zeros((n_meas, n_t_traj, 3), np.float32)
        for m in range(n_meas):
            for c in range(3):
                Gt[m, :, c] = np.interp(t_tr, t_wf, G[m, :, c], left=0.0, right=0.0)
    elif n_t_wf >= n_t_traj:
        Gt = G[:, :n_t_traj, :]
    else:
        Gt = np.zeros((n_meas, n_t_traj, 3), np.float32)
        Gt[:, :n_t_wf, :] = G
    chi = np.ones(n_t_wf) if chi is None else np.asarray(chi, np.float64)
    step = t_wf[-1] / max(n_t_wf - 1, 1) if n_t_wf > 1 else dt_traj
    idx = np.round(t_tr / step).astype(int)
    ok = (idx >= 0) & (idx < n_t_wf)
    cl = np.clip(idx, 0, n_t_wf - 1)
    if chi.ndim == 2:
        chi_r = np.where(ok[None, :], chi[:, cl], 0.0)
    else:
        chi_r = np.where(ok, chi[cl], 0.0)[None, :]
    return Gt, chi_r


def replay(traj, dt_traj, G, dt_wf, chi=None, dlog=None, T2=None, T1=None, rho=None, D=None,
           stimulated_echo=False):
    """(signals (n_meas,), phi (n_meas, N), log_w (n_meas, N or 1)) in float64."""
    traj = np.asarray(traj)
    N, n_t, _ = traj.shape
    Gt, chi_r = resample(G, dt_wf, dt_traj, n_t, chi)
    n_meas = Gt.shape[0]
    phi = GAMMA * dt_traj * (Gt.reshape(n_meas, 3 * n_t).astype(np.float64)
                             @ traj.reshape(N, 3 * n_t).astype(np.float64).T)
    lw = np.zeros((n_meas, 1))
    if T2 is not None:
        lw[:, 0] -= (dt_traj / T2) * chi_r.sum(axis=1)
    if T1 is not None:
        lw[:, 0] -= (dt_traj / T1) * (n_t - chi_r.sum(axis=1))
    if rho is not None:
        lw = lw + (rho / D) * (chi_r @ np.asarray(dlog, np.float64).T)
    S = np.mean(np.exp(lw) * np.cos(phi), axis=1)
    if stimulated_echo:
        S = S * 0.5
    return S, phi, lw
