"""Walk once, replay many waveforms (dmipy_sim/core.py:972-1529,
dmipy_sim/trajectories.py:360-641).

``simulate_trajectories`` walks the spins with no gradient and keeps float16
positions after every saved sample (plus, on request, the boundary log-weight
at rho/D = 1 and the compartment occupancy of permeable walls).
``replay`` applies a waveform and relaxation to saved trajectories:

    phi[m,w]   = GAMMA dt_traj sum_t G[m,t,:] . r[w,t,:]                (float64)
    log_w[m,w] = -dt/T2 sum_t chi[m,t] - dt/T1 sum_t (1 - chi[m,t])
                 + (rho/D) sum_t chi[m,t] dlog[w,t]
    S[m]       = mean_w exp(log_w[m,w]) cos(phi[m,w])

Both halves run on the GPU: the producer is the walk kernel writing float16
channels, the contraction is a float64 matrix-core kernel (csrc/replay.hip).
``TrajectorySet`` keeps the trajectories in device memory so any number of
waveforms can be replayed without moving them; the module-level functions keep
the reference's signatures and return numpy arrays.
"""
import ctypes as C
import math

import warnings

import numpy as np

from . import _native as nat
from . import geometries as geo
from .constants import GAMMA
from .physics import make_walk_plan, WALL_REFLECT, WALL_SURFACE, WALL_PERMEATE

f32 = np.float32


def producer_radius(geometry):
    """Length scale of the producer's sub-step rule: radius, sphere radius,
    Box1D length, smallest packing radius, smallest inner radius
    (dmipy_sim/core.py:1088-1106); None for free diffusion."""
    R = getattr(geometry, 'radius', None)
    if R is None:
        R = getattr(geometry, 'sphere_radius', None)
    if R is None:
        R = getattr(geometry, 'length', None)
    if R is None:
        radii = getattr(geometry, '_radii_np', None)
        if radii is not None and len(radii) > 0:
            R = float(np.min(radii))
    if R is None:
        inner = getattr(geometry, '_inner_radii_np', None)
        if inner is not None and len(inner) > 0 and np.any(inner > 0):
            R = float(np.min(inner[inner > 0]))
    return None if R is None else float(R)


def producer_plan(geometry, diffusivity, T_max, dt_save, save_relaxation_data=False):
    """Sample grid, sub-steps and step constants of the producer walk
    (dmipy_sim/core.py:1084-1148): n_t = round(T_max/dt_save) + 1, dt =
    T_max/(n_t-1), sub_steps = ceil(dt / (R^2/(216 D))) (3750 when permeable),
    step = float32 sqrt of float32(6 D dt_sim); the wall records its log-weight
    at rho/D = 1 when relaxation data are saved."""
    n_t = int(round(T_max / dt_save)) + 1
    if n_t < 2:
        raise ValueError("T_max must span at least one saved step")
    dt_actual = T_max / (n_t - 1)
    R = producer_radius(geometry)
    has_perm = getattr(geometry, 'permeability', None) is not None
    if R is None:
        sub_steps = 1
    else:
        dt_max = R ** 2 / ((3750.0 if has_perm else 216.0) * diffusivity)
        sub_steps = max(1, int(np.ceil(dt_actual / dt_max)))
    dt_sim = dt_actual / sub_steps
    step_len = float(np.sqrt(f32(6.0 * diffusivity * dt_sim), dtype=np.float32))
    plan = make_walk_plan(geometry, diffusivity, dt_actual)
    kind = plan.kind
    if kind == geo.KIND_FREE:
        wall, rhoD = WALL_REFLECT, 0.0
    elif has_perm:
        wall, rhoD = WALL_PERMEATE, (1.0 if save_relaxation_data else 0.0)
    elif save_relaxation_data:
        wall, rhoD = WALL_SURFACE, 1.0
    else:
        wall, rhoD = WALL_REFLECT, 0.0
    kappaD = float(f32(float(geometry.permeability) / diffusivity)) if has_perm else 0.0
    return dict(n_t=n_t, dt_actual=dt_actual, sub_steps=sub_steps, dt_sim=dt_sim,
                step_len=step_len, wall=wall, rho_over_D=rhoD, kappa_over_D=kappaD,
                plan=plan, permeable=has_perm)


# Geometries whose producer records compartment occupancy when permeable: the
# reference's rule keys on a `radius` attribute (dmipy_sim/core.py:1166-1185).
_OCCUPANCY_KINDS = (geo.KIND_CYLINDER, geo.KIND_SPHERE, geo.KIND_SLAB, geo.KIND_SHELL,
                    geo.KIND_SHELL_CYL)


def _half(a):
    return np.ascontiguousarray(a, dtype=np.float16).view(np.uint16)


class TrajectorySet:
    """Trajectories resident in device memory (float16, one walk)."""

    def __init__(self, handle, dt_traj, sub_steps=1, dt_sim=None, comp_kind=None):
        self._h = handle
        nat.track(self)
        n = C.c_int64()
        t = C.c_int32()
        ch = C.c_int32()
        ms = C.c_double()
        nat.check(nat.lib().dmw_traj_info(handle, C.byref(n), C.byref(t), C.byref(ch), C.byref(ms)))
        self.n_walkers, self.n_t, self.channels = int(n.value), int(t.value), int(ch.value)
        self.walk_ms = float(ms.value)
        self.dt_traj = float(dt_traj)
        self.sub_steps = int(sub_steps)
        self.dt_sim = float(dt_sim) if dt_sim is not None else self.dt_traj
        self.comp_kind = comp_kind          # None, 'occupancy', 'zeros' or 'zeros16'

    @classmethod
    def simulate(cls, n_walkers, diffusivity, geometry, T_max, dt_save, seed=42,
                 save_relaxation_data=False, r0=None):
        """The producer walk (dmipy_sim/core.py:972-1529) on the GPU."""
        pp = producer_plan(geometry, diffusivity, T_max, dt_save, save_relaxation_data)
        plan = pp["plan"]
        N = int(n_walkers)
        if N < 1:
            raise ValueError("n_walkers must be >= 1")
        rot = plan.rot
        r0n = None
        if r0 is not None:
            r0n = np.asarray(r0, np.float64).reshape(N, 3)
            if rot is not None:
                r0n = (rot.astype(np.float64) @ r0n.T).T
            r0n = np.ascontiguousarray(r0n, dtype=np.float32)
        gd = plan.geometry_desc(wall=pp["wall"], rho_over_D=pp["rho_over_D"],
                                  kappa_over_D=pp["kappa_over_D"])
        rd = nat.Run_t(n_walkers=N, seed=int(seed) & (2**64 - 1), n_sub=pp["sub_steps"],
                       step_len=pp["step_len"], gamma_dt_sub=0.0, dt_sub=float(f32(pp["dt_sim"])),
                       inv_T2=0.0, inv_T1=0.0, has_t2=0, has_t1=0, kahan=0, meas_block=0,
                       r0=nat.fptr(r0n), inj_normals=None, inj_uniforms=None)
        comp_kind = None
        channels = 0
        if save_relaxation_data:
            channels |= 1
            if pp["permeable"] and plan.kind in _OCCUPANCY_KINDS:
                channels |= 2
                comp_kind = "occupancy"
            else:
                # permeable without a radius (Ellipsoid): float16 zeros; else int8
                comp_kind = "zeros16" if pp["permeable"] else "zeros"
        lab = None
        if rot is not None:
            lab = np.ascontiguousarray(np.asarray(rot, np.float64).T, dtype=np.float32)
        h = C.c_void_p()
        nat.check(nat.lib().dmw_traj_create(C.byref(h), gd, rd, pp["n_t"], channels,
                                            nat.fptr(lab)))
        return cls(h, pp["dt_actual"], pp["sub_steps"], pp["dt_sim"], comp_kind)

    @classmethod
    def from_arrays(cls, trajectory, dt_traj, dlog_boundary_unit=None):
        """Upload (n_walkers, n_t, 3) positions and optional (n_walkers, n_t)
        boundary log-weights (stored as float16, as the producer saves them)."""
        tr = np.asarray(trajectory)
        if tr.ndim != 3 or tr.shape[2] != 3:
            raise ValueError("trajectory must have shape (n_walkers, n_t, 3)")
        N, n_t, _ = tr.shape
        th = _half(tr)
        dh = None
        if dlog_boundary_unit is not None:
            d = np.asarray(dlog_boundary_unit)
            if d.shape != (N, n_t):
                raise ValueError("dlog_boundary_unit must have shape (n_walkers, n_t)")
            dh = _half(d)
        h = C.c_void_p()
        u16 = C.POINTER(C.c_uint16)
        nat.check(nat.lib().dmw_traj_from_host(C.byref(h), N, n_t, th.ctypes.data_as(u16),
                                               None if dh is None else dh.ctypes.data_as(u16)))
        return cls(h, dt_traj)

    def download(self):
        """(trajectories (N, n_t, 3) float16, dlog (N, n_t) float16 or None,
        comp (N, n_t) float16 / int8 or None), as simulate_trajectories returns them."""
        u16 = C.POINTER(C.c_uint16)
        tr = np.empty((self.n_walkers, self.n_t, 3), np.uint16)
        dl = np.empty((self.n_walkers, self.n_t), np.uint16) if self.channels & 1 else None
        cp = np.empty((self.n_walkers, self.n_t), np.uint16) if self.channels & 2 else None
        nat.check(nat.lib().dmw_traj_download(
            self._h, tr.ctypes.data_as(u16), None if dl is None else dl.ctypes.data_as(u16),
            None if cp is None else cp.ctypes.data_as(u16)))
        comp = None
        if cp is not None:
            comp = cp.view(np.float16)
        elif self.comp_kind in ("zeros", "zeros16"):
            comp = np.zeros((self.n_walkers, self.n_t),
                            np.float16 if self.comp_kind == "zeros16" else np.int8)
        return tr.view(np.float16), (None if dl is None else dl.view(np.float16)), comp

    def replay(self, G, dt_wf, *, chi_perp=None, T2=None, T1=None, surface_relaxivity=None,
               D=None, stimulated_echo=False, return_walker_signals=False, meas_tiles=0):
        """Apply a waveform (dmipy_sim/trajectories.py:360-641) to these
        trajectories; returns signals (n_meas,) float64, or (phi, log_w_total,
        signals) with ``return_walker_signals``."""
        G_traj, chi_r, scalar, per_meas = _prepare(G, dt_wf, self.dt_traj, self.n_t, chi_perp, T2, T1)
        n_meas = G_traj.shape[0]
        rhoD = 0.0
        if surface_relaxivity is not None:
            if D is None:
                raise ValueError(
                    "D (diffusivity) must be provided when surface_relaxivity is not None.")
            if not self.channels & 1:
                raise ValueError(
                    "dlog_boundary_unit must be provided when surface_relaxivity is not None.  "
                    "Re-run simulate_trajectories with save_relaxation_data=True.")
            rhoD = float(surface_relaxivity) / float(D)
        Gd = np.ascontiguousarray(G_traj, dtype=np.float64)
        chi = np.ascontiguousarray(chi_r if per_meas else chi_r[0], dtype=np.float64)
        lw0 = np.ascontiguousarray(scalar, dtype=np.float64)
        rp = nat.Replay_t(n_meas=n_meas, n_t=self.n_t, G=nat.dptr(Gd),
                          scale=GAMMA * self.dt_traj, chi=nat.dptr(chi),
                          chi_rows=n_meas if per_meas else 1, log_w0=nat.dptr(lw0),
                          rho_over_D=rhoD, meas_tiles=int(meas_tiles))
        sig = np.zeros(n_meas, np.float64)
        phi = lw = None
        if return_walker_signals:
            phi = np.zeros((n_meas, self.n_walkers), np.float64)
            lw = np.zeros((n_meas, self.n_walkers), np.float64)
        ms = C.c_double()
        nat.check(nat.lib().dmw_replay_run(self._h, rp, nat.dptr(sig), nat.dptr(phi), nat.dptr(lw),
                                           C.byref(ms)))
        self.last_replay_ms = float(ms.value)
        if stimulated_echo:
            sig = sig * 0.5
        if return_walker_signals:
            if surface_relaxivity is None:
                lw = lw[:, :1].copy()          # walker-independent: (n_meas, 1)
            return phi, lw, sig
        return sig

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            nat.lib().dmw_traj_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _prepare(G, dt_wf, dt_traj, n_t_traj, chi_perp, T2, T1):
    """Waveform and gate on the trajectory grid plus the walker-independent
    log-weight, as dmipy_sim/trajectories.py:501-591 forms them: linear
    resampling of G (zero outside the waveform window) when the grids differ,
    truncation or zero padding when only the lengths differ; nearest-sample
    chi (zero past the waveform end); T2 and T1 sums over the gate."""
    G = np.asarray(G, dtype=np.float32)
    n_meas, n_t_wf, _ = G.shape
    if chi_perp is None:
        chi_perp = np.ones(n_t_wf)
    chi_perp = np.asarray(chi_perp, dtype=np.float64)
    per_meas = chi_perp.ndim == 2
    rel = abs(dt_wf - dt_traj) / max(abs(dt_traj), 1e-30)
    if rel > 1e-9:
        t_wf = np.arange(n_t_wf, dtype=np.float64) * dt_wf
        t_tr = np.arange(n_t_traj, dtype=np.float64) * dt_traj
        G_traj = np.zeros((n_meas, n_t_traj, 3), dtype=np.float32)
        for m in range(n_meas):
            for ax in range(3):
                G_traj[m, :, ax] = np.interp(t_tr, t_wf, G[m, :, ax], left=0.0, right=0.0)
    elif n_t_wf == n_t_traj:
        G_traj = G
    elif n_t_wf > n_t_traj:
        G_traj = G[:, :n_t_traj, :]
    else:
        G_traj = np.zeros((n_meas, n_t_traj, 3), dtype=np.float32)
        G_traj[:, :n_t_wf, :] = G
    t_wf = np.arange(n_t_wf, dtype=np.float64) * dt_wf
    t_tr = np.arange(n_t_traj, dtype=np.float64) * dt_traj

    def resample(c):
        dt_eff = t_wf[-1] / max(n_t_wf - 1, 1) if n_t_wf > 1 else dt_traj
        idx = np.round(t_tr / dt_eff).astype(int)
        inside = (idx >= 0) & (idx < n_t_wf)
        return np.where(inside, c[np.clip(idx, 0, n_t_wf - 1)], 0.0)

    chi_r = (np.stack([resample(chi_perp[m]) for m in range(n_meas)]) if per_meas
             else resample(chi_perp)[np.newaxis, :])
    scalar = np.zeros(n_meas, dtype=np.float64)
    if T2 is not None:
        scalar -= (dt_traj / T2) * chi_r.sum(axis=1)
    if T1 is not None:
        scalar -= (dt_traj / T1) * (n_t_traj - chi_r.sum(axis=1))
    return G_traj, chi_r, scalar, per_meas


def unwrap_periodic(traj, cell_size, periodic_axes=(0, 1)):
    """Continuous positions from positions folded into a periodic cell
    (dmipy_sim/trajectories.py:82-132): along each periodic axis the per-save
    displacement is taken to its minimal image, x_k - x_{k-1} - L round((x_k -
    x_{k-1}) / L), and re-integrated from the first sample.  Time is axis -2 of
    ``traj`` (..., n_t, 3); valid while the true per-save displacement stays below
    L/2 (a warning is raised when one comes within 0.45 L).  float32.

    The walk-once producer of this library already saves unfolded positions, so
    this is for positions from elsewhere (e.g. a folded export)."""
    out = np.array(traj, dtype=np.float32, copy=True)
    L = np.float32(cell_size)
    for ax in periodic_axes:
        x = out[..., ax]
        step = np.diff(x, axis=-1)
        step = step - L * np.round(step / L)
        if np.any(np.abs(step) > 0.45 * float(L)):
            warnings.warn("unwrap_periodic: a per-save displacement approaches L/2 on a periodic "
                          "axis; the save step may be too coarse to unwrap unambiguously.",
                          RuntimeWarning, stacklevel=2)
        cont = np.empty_like(x)
        cont[..., 0] = x[..., 0]
        cont[..., 1:] = x[..., :1] + np.cumsum(step, axis=-1)
        out[..., ax] = cont
    return out


def simulate_trajectories(n_walkers, diffusivity, geometry, T_max, dt_save, seed=42,
                          walker_batch_size=50_000, save_relaxation_data=False, require_gpu=None,
                          r0=None, kappa_MT=0.0, dwell_time=0.0, equilibrate_binding="auto",
                          compress=None):
    """Walk once and save float16 positions after every sample
    (dmipy_sim/core.py:972-1529).  Returns (trajectories (N, n_t, 3) float16,
    dt_actual, sub_steps, dt_sim[, dlog_boundary_unit (N, n_t) float16,
    comp_traj (N, n_t)]).  The whole walker set is walked in one launch
    (``walker_batch_size`` only bounds the reference's host batches)."""
    from .gpu import check_gpu
    if kappa_MT:
        raise NotImplementedError("magnetization-transfer binding belongs to the packed "
                                  "myelin geometry, which this build does not provide")
    if compress:
        raise NotImplementedError("compressed (temporal DCT) trajectory masters are not provided")
    check_gpu(n_walkers, require_gpu, what="simulate_trajectories")
    ts = TrajectorySet.simulate(n_walkers, diffusivity, geometry, T_max, dt_save, seed=seed,
                                save_relaxation_data=save_relaxation_data, r0=r0)
    try:
        traj, dlog, comp = ts.download()
    finally:
        ts.close()
    if save_relaxation_data:
        return traj, ts.dt_traj, ts.sub_steps, ts.dt_sim, dlog, comp
    return traj, ts.dt_traj, ts.sub_steps, ts.dt_sim


def replay(trajectory, dt_traj, G, dt_wf, *, chi_perp=None, dlog_boundary_unit=None, T2=None,
           T1=None, surface_relaxivity=None, D=None, stimulated_echo=False, comp_traj=None,
           T2_per_comp=None, T1_per_comp=None, return_walker_signals=False, susceptibility=None,
           eps_P=None):
    """Apply a gradient waveform and relaxation weights to saved trajectories
    (dmipy_sim/trajectories.py:360-641).  ``trajectory`` is an (N, n_t, 3)
    array (float16 or float32; stored as float16) or a TrajectorySet."""
    if isinstance(trajectory, dict):
        raise NotImplementedError("compressed trajectory masters are not provided")
    if T2_per_comp is not None or T1_per_comp is not None:
        if comp_traj is None:
            raise ValueError("comp_traj required when T2_per_comp is provided."
                             if T2_per_comp is not None else
                             "comp_traj required when T1_per_comp is provided.")
        raise NotImplementedError("per-compartment relaxation replay serves the Mesh geometry, "
                                  "which this build does not provide")
    if susceptibility is not None:
        raise NotImplementedError("susceptibility off-resonance replay is not provided")
    if surface_relaxivity is not None:
        if D is None:
            raise ValueError(
                "D (diffusivity) must be provided when surface_relaxivity is not None.")
        if dlog_boundary_unit is None and not isinstance(trajectory, TrajectorySet):
            raise ValueError(
                "dlog_boundary_unit must be provided when surface_relaxivity is not None.  "
                "Re-run simulate_trajectories with save_relaxation_data=True.")
    if isinstance(trajectory, TrajectorySet):
        ts, own = trajectory, False
    else:
        ts, own = TrajectorySet.from_arrays(trajectory, dt_traj, dlog_boundary_unit), True
        ts.dt_traj = float(dt_traj)
    try:
        return ts.replay(G, dt_wf, chi_perp=chi_perp, T2=T2, T1=T1,
                         surface_relaxivity=surface_relaxivity, D=D,
                         stimulated_echo=stimulated_echo,
                         return_walker_signals=return_walker_signals)
    finally:
        if own:
            ts.close()
