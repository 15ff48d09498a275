"""Many substrate instances in one persistent, graph-captured launch.

Each instance is any geometry the walk supports (free, Box1D, Cylinder, Sphere,
PackedCylinders, PackedSpheres; reflecting, relaxing or permeable walls) with
its own diffusivity, T2, walker count and sub-step count, and names one of a
list of waveforms (all with the same number of measurements, <= 64 -- up to 128
when every instance is a permeable packing, whose kernel parks the phase in
device memory -- and the same n_t).  The result for instance i equals, in distribution, a separate
``simulate(n_walkers_i, D_i, waveforms[k_i], geometry_i)`` call
(dmipy_sim/core.py:193) -- the serial per-substrate loop the reference runs,
e.g. simulate_mixture (core.py:821-858).  By default instance i uses the Philox
instance word i, so instance 0 walks exactly as ``walk(..., seed)`` does; with
``stream=0, seed_offset=k`` an instance walks as ``walk(..., seed + k)``.
"""
import ctypes as C

import numpy as np

from . import _native as nat
from .physics import make_walk_plan


def _screened(geometry):
    """True for the permeable packings, which walk in the screened kernel."""
    from . import geometries as geo
    return (isinstance(geometry, (geo.PackedCylinders, geo.PackedSpheres))
            and getattr(geometry, "permeability", None) is not None)


class BatchedWalk:
    """Plan once (upload, cost-sorted work queue, hipGraph), execute per seed.

    ``instances`` is a list of dicts with keys 'geometry', 'diffusivity',
    'n_walkers' and optionally 'T2', 'waveform' (index into ``waveforms``,
    default 0), 'stream' (Philox instance word, default the instance index) and
    'seed_offset' (added to the run seed, default 0).  ``waveforms`` is one
    Waveform or a list of them.

    Launch shape (results do not depend on it): ``waves_per_simd`` persistent
    wavefronts per SIMD (0: 4), ``walkers_per_wave`` walkers a wavefront takes
    per queue pop (64, 128, 256, 512; 0: 64), ``unroll`` of the sub-step loop
    (1, 2; 0: 1).  ``tuning="auto"`` takes the autotuner's cached choice for
    this batch's key (see autotune.tune_batch), or the defaults if none.
    """

    def __init__(self, instances, waveforms, waves_per_simd=0, walkers_per_wave=0, unroll=0,
                 use_graph=True, tuning=None):
        if not isinstance(waveforms, (list, tuple)):
            waveforms = [waveforms]
        if len(instances) < 1 or len(waveforms) < 1:
            raise ValueError("need at least one instance and one waveform")
        Gs = [np.asarray(w.G, dtype=np.float32) for w in waveforms]
        n_meas, n_t, _ = Gs[0].shape
        for G in Gs:
            if G.shape != (n_meas, n_t, 3):
                raise ValueError("every waveform needs the same (n_measurements, n_t)")
        if n_meas > 128:
            raise ValueError("batched walks take at most 128 measurements per waveform")
        if n_meas > 64 and not all(_screened(s["geometry"]) for s in instances):
            raise ValueError("batched walks of more than 64 measurements need every instance "
                             "to be a permeable packing")
        chis = [getattr(w, 'chi_perp', None) for w in waveforms]
        chi0 = None if chis[0] is None else np.asarray(chis[0], np.float32).reshape(n_t)
        for c in chis[1:]:
            same = (c is None and chi0 is None) or (
                c is not None and chi0 is not None and np.array_equal(
                    np.asarray(c, np.float32).reshape(n_t), chi0))
            if not same:
                raise ValueError("every waveform of a batch needs the same chi_perp")
        dts = {float(w.dt) for w in waveforms}

        arr = (nat.Instance_t * len(instances))()
        self.n_sub = np.zeros(len(instances), np.int64)
        self.n_walkers = np.zeros(len(instances), np.int64)
        self._half = np.ones(len(instances), np.float64)
        set_key = {}
        sets_G = []
        kind_walls = set()
        self._keep = []          # host arrays referenced by the descriptors
        for i, spec in enumerate(instances):
            k = int(spec.get('waveform', 0))
            if not 0 <= k < len(waveforms):
                raise ValueError(f"instance {i}: waveform index {k} out of range")
            wf = waveforms[k]
            if spec.get('T1') is not None:
                raise NotImplementedError("batched walks do not model T1")
            plan = make_walk_plan(spec['geometry'], spec['diffusivity'], wf.dt, T2=spec.get('T2'))
            rot = plan.rot
            key = (k, None if rot is None else np.asarray(rot, np.float64).tobytes())
            if key not in set_key:
                G = Gs[k]
                if rot is not None:
                    G = np.einsum('ij,mtj->mti', np.asarray(rot, np.float64), G)
                set_key[key] = len(sets_G)
                sets_G.append(np.ascontiguousarray(G, dtype=np.float32))
            gd = plan.geometry_desc()
            self._keep.append(plan)      # owns the arrays the descriptor points into
            arr[i] = nat.Instance_t(geom=gd, n_walkers=int(spec['n_walkers']), n_sub=plan.n_sub,
                                    step_len=plan.step_len, gamma_dt_sub=plan.gamma_dt_sub,
                                    dt_sub=plan.dt_sub, inv_T2=plan.inv_T2,
                                    has_t2=int(plan.has_t2), wave_set=set_key[key],
                                    stream=int(spec.get('stream', -1)),
                                    seed_offset=int(spec.get('seed_offset', 0)) & (2**64 - 1))
            self.n_sub[i] = plan.n_sub
            self.n_walkers[i] = int(spec['n_walkers'])
            kind_walls.add((gd.kind, gd.wall))
            if getattr(wf, 'stimulated_echo', False):
                self._half[i] = 0.5
        self._chi = None if chi0 is None else np.ascontiguousarray(chi0)
        self._sets = sets_G
        wds = (nat.Waveform_t * len(sets_G))()
        for s, G in enumerate(sets_G):
            wds[s] = nat.Waveform_t(n_meas=n_meas, n_t=n_t, G=nat.fptr(G),
                                    chi=nat.fptr(self._chi), n_echo=0, echo_idx=None)
        from . import autotune
        self.key = autotune.batch_key(kind_walls, self.n_sub, n_meas)
        if tuning == "auto":
            waves_per_simd, walkers_per_wave, unroll = autotune.cached_batch_shape(self.key)
        elif tuning is not None:
            raise ValueError("tuning must be None or 'auto'")
        self.shape = dict(waves_per_simd=int(waves_per_simd), walkers_per_wave=int(walkers_per_wave),
                          unroll=int(unroll))
        tn = nat.BatchTuning_t(**self.shape)
        h = C.c_void_p()
        nat.check(nat.lib().dmw_batch_create(C.byref(h), len(instances), arr, len(sets_G), wds,
                                             C.byref(tn), int(bool(use_graph))))
        self._h = h
        self._keep = None
        nat.track(self)
        self.n_inst = len(instances)
        self.n_sets = len(sets_G)
        self.n_meas = n_meas
        self.n_t = n_t
        self.dt = dts.pop() if len(dts) == 1 else None

    def run(self, seed=123):
        """Signals (n_instances, n_measurements) float64 and timing stats."""
        out = np.zeros((self.n_inst, self.n_meas), np.float64)
        st = nat.BatchStats_t()
        nat.check(nat.lib().dmw_batch_execute(self._h, int(seed) & (2**64 - 1),
                                              nat.dptr(out), C.byref(st)))
        out *= self._half[:, None]
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
