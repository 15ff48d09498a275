"""Many substrate instances in one persistent, graph-captured launch.

Each instance is a ``Cylinder`` or ``Sphere`` (reflecting, relaxing or permeable
wall) with its own diffusivity, walker count and sub-step count; all share one
waveform.  The result for instance i equals, in distribution, a separate
``simulate(n_walkers_i, D_i, waveform, geometry_i)`` call (dmipy_sim/core.py:193),
i.e. the serial per-substrate loop the reference would run.
"""
import ctypes as C

import numpy as np

from . import _native as nat
from . import geometries as geo
from .physics import make_walk_plan


class BatchedWalk:
    """Plan once (upload, cost-sorted work queue, hipGraph), execute per seed.

    ``instances`` is a list of dicts with keys 'geometry', 'diffusivity',
    'n_walkers' and optionally 'T2'.
    """

    def __init__(self, instances, waveform, waves_per_simd=0, use_graph=True):
        G = np.ascontiguousarray(np.asarray(waveform.G, dtype=np.float32))
        n_meas, n_t, _ = G.shape
        arr = (nat.Instance_t * len(instances))()
        self.n_sub = np.zeros(len(instances), np.int64)
        self.n_walkers = np.zeros(len(instances), np.int64)
        for i, spec in enumerate(instances):
            g = spec['geometry']
            if g.kind not in (geo.KIND_CYLINDER, geo.KIND_SPHERE):
                raise NotImplementedError("batched instances must be Cylinder or Sphere")
            if getattr(g, '_R', None) is not None and not g._is_identity_rotation:
                raise NotImplementedError("batched cylinders must be along z")
            plan = make_walk_plan(g, spec['diffusivity'], waveform.dt, T2=spec.get('T2'))
            gd = nat.Geometry_t(kind=plan.kind, wall=plan.wall, n_obj=0, size=plan.size,
                                cell=0.0, eps=plan.eps, nudge=plan.nudge,
                                rho_over_D=plan.rho_over_D, kappa_over_D=plan.kappa_over_D,
                                centers=None, radii=None)
            arr[i] = nat.Instance_t(geom=gd, n_walkers=int(spec['n_walkers']), n_sub=plan.n_sub,
                                    step_len=plan.step_len, gamma_dt_sub=plan.gamma_dt_sub,
                                    dt_sub=plan.dt_sub, inv_T2=plan.inv_T2,
                                    has_t2=int(plan.has_t2))
            self.n_sub[i] = plan.n_sub
            self.n_walkers[i] = int(spec['n_walkers'])
        chi = getattr(waveform, 'chi_perp', None)
        self._chi = None if chi is None else np.ascontiguousarray(chi, dtype=np.float32).reshape(n_t)
        self._G = G
        wd = nat.Waveform_t(n_meas=n_meas, n_t=n_t, G=nat.fptr(G), chi=nat.fptr(self._chi),
                            n_echo=0, echo_idx=None)
        h = C.c_void_p()
        nat.check(nat.lib().dmw_batch_create(C.byref(h), len(instances), arr, wd,
                                             int(waves_per_simd), int(bool(use_graph))))
        self._h = h
        self.n_inst = len(instances)
        self.n_meas = n_meas
        self.n_t = n_t
        self.stimulated_echo = bool(getattr(waveform, 'stimulated_echo', False))

    def run(self, seed=123):
        """Signals (n_instances, n_measurements) float64 and timing stats."""
        out = np.zeros((self.n_inst, self.n_meas), np.float64)
        st = nat.BatchStats_t()
        nat.check(nat.lib().dmw_batch_execute(self._h, int(seed) & (2**64 - 1),
                                              nat.dptr(out), C.byref(st)))
        if self.stimulated_echo:
            out *= 0.5
        return out, {"kernel_ms": st.kernel_ms, "total_ms": st.total_ms,
                     "walker_steps": int(st.walker_steps), "n_units": st.n_units,
                     "grid": st.grid}

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            nat.lib().dmw_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
