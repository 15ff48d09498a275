"""Sub-step rules and the per-run constants of one walk.

``make_walk_plan`` is the counterpart of ``make_step_fn``
(dmipy_sim/physics.py:191-417): it decides which wall routine the walk uses
(plain reflection, reflection with the Brownstein-Tarr weight, or Powles
permeation), how many physical sub-steps resolve one waveform sample, and the
float32 constants of the step.  Each constant is rounded exactly where the
reference rounds it, so the kernel and a restatement of the reference step agree
bit for bit on identical increments:

* step length  ``sqrt(f32(6*D) * f32(dt/n_sub))``      (physics.py:256-260)
* gamma*dt_sub ``f32(GAMMA*dt/n_sub)``                  (physics.py:288-290)
* rho/D        ``f32(rho)/f32(D)`` on the surface branch (physics.py:322-328),
               ``f32(rho/D)`` on the permeable branch    (physics.py:279-281)
* kappa/D      ``f32(kappa/D)``                         (physics.py:279)
* EPS, NUDGE   ``f32(1e-7*R)``, ``f32(1e-4*R)`` with R the radius, or the
               smallest radius of a packing (geometries.py:503-504, 1694-1696)
"""
import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from .constants import GAMMA
from . import geometries as geo

WALL_REFLECT, WALL_SURFACE, WALL_PERMEATE = 0, 1, 2

f32 = np.float32


def _geometry_radius(geometry):
    """Smallest geometric radius, or None (dmipy_sim/physics.py:15-32)."""
    R = getattr(geometry, 'radius', None)
    if R is None:
        R = getattr(geometry, 'sphere_radius', None)
    if R is None:
        radii = getattr(geometry, '_radii_np', None)
        if radii is not None and len(radii) > 0:
            R = float(np.min(radii))
    if R is None:
        inner = getattr(geometry, '_inner_radii_np', None)
        if inner is not None and len(inner) > 0 and np.any(inner > 0):
            R = float(np.min(inner[inner > 0]))
    return float(R) if R is not None else None


def permeable_sub_steps(geometry, diffusivity, dt):
    """Sub-steps so the step is about R/25 (dt_max = R^2/(3750 D)); 1 without a
    radius (dmipy_sim/physics.py:35-48)."""
    R = _geometry_radius(geometry)
    if R is None:
        return 1
    dt_max = R ** 2 / (3750.0 * diffusivity)
    return max(1, int(np.ceil(dt / dt_max)))


def _surface_char_radius(geometry):
    """Pore scale for the surface rule: the extra-axonal pore area/perimeter when
    the geometry describes one, else the geometric radius
    (dmipy_sim/physics.py:51-69)."""
    outer = getattr(geometry, '_outer_radii_np', None)
    cell = getattr(geometry, '_cell_size', None)
    if outer is not None and cell is not None:
        outer = np.asarray(outer, float)
        outer = outer[outer > 0]
        area_ext = float(cell) ** 2 - float(np.sum(np.pi * outer ** 2))
        perim = float(np.sum(2.0 * np.pi * outer))
        if perim > 0 and area_ext > 0:
            return area_ext / perim
    return _geometry_radius(geometry)


def surface_sub_steps(geometry, diffusivity, dt, frac=8.0):
    """Sub-steps so the step is about R_char/frac; ``surface_substep_frac`` on the
    geometry overrides ``frac`` and 0/None disables sub-stepping
    (dmipy_sim/physics.py:163-188)."""
    g_frac = getattr(geometry, 'surface_substep_frac', None)
    if g_frac is not None:
        frac = g_frac
    if not frac or frac <= 0:
        return 1
    Rc = _surface_char_radius(geometry)
    if Rc is None:
        return 1
    dt_max = (Rc / frac) ** 2 / (6.0 * diffusivity)
    return max(1, int(np.ceil(dt / dt_max)))


def collision_sub_steps(geometry, diffusivity, dt, frac=0.9):
    """Sub-steps keeping a step inside a collision grid cell; 1 for geometries
    without ``cell_size`` -- every analytic geometry here
    (dmipy_sim/physics.py:72-95)."""
    cs = getattr(geometry, "cell_size", None)
    if cs is None or not np.isfinite(cs) or cs <= 0:
        return 1
    L = float(np.sqrt(6.0 * float(diffusivity) * float(dt)))
    ratio = L / (float(frac) * float(cs))
    return int(max(1, np.ceil(ratio ** 2))) if ratio > 1.0 else 1


def replay_sub_steps(geometry, diffusivity, dt, permeable=False):
    """Sub-step rule of the walk-once producer: step <= R/6 (dt_max = R^2/(216 D)),
    R/25 when permeable (dmipy_sim/core.py:1089-1118)."""
    R = _geometry_radius(geometry)
    if R is None:
        return 1
    div = 3750.0 if permeable else 216.0
    return max(1, int(np.ceil(dt / (R ** 2 / (div * diffusivity)))))


@dataclass
class WalkPlan:
    """Everything the kernel needs besides the waveform and the walkers."""
    kind: int
    wall: int
    n_sub: int
    size: float            # radius (Cylinder/Sphere), length (Box1D)
    cell: float            # periodic cell side (packed)
    eps: float
    nudge: float
    rho_over_D: float
    kappa_over_D: float
    step_len: float
    gamma_dt_sub: float
    dt_sub: float
    inv_T2: float
    inv_T1: float
    has_t2: bool
    has_t1: bool
    has_weight: bool
    centers: np.ndarray = None   # native frame, float32
    radii: np.ndarray = None     # float32
    rot: np.ndarray = None       # lab -> native rotation (float32 3x3) or None
    extra: dict = field(default_factory=dict)
    axes: tuple = (0.0, 0.0, 0.0)  # Ellipsoid: float32 1/a^2, 1/b^2, 1/c^2
    aux: float = 0.0               # Ellipsoid clamp factor; shell / myelin outer radius
    myelin: dict = None            # MyelinatedCylinder per-compartment constants
    packed_myelin: dict = None     # PackedMyelinatedCylinders per-axon tables

    def geometry_desc(self, wall=None, rho_over_D=None, kappa_over_D=None, n_walkers=None):
        """The dmw_geometry descriptor of this plan (optionally with another wall
        mode or wall constants, as the walk-once producer uses).  The myelinated
        cylinder's start allotment depends on ``n_walkers``."""
        from . import _native as nat
        c = None if self.centers is None else np.ascontiguousarray(self.centers, np.float32)
        r = None if self.radii is None else np.ascontiguousarray(self.radii, np.float32)
        self.extra["_desc_keep"] = (c, r)          # keep the arrays alive with the plan
        g = nat.Geometry_t(kind=self.kind, wall=self.wall if wall is None else wall,
                           n_obj=0 if r is None else len(r), size=self.size, cell=self.cell,
                           eps=self.eps, nudge=self.nudge,
                           rho_over_D=self.rho_over_D if rho_over_D is None else rho_over_D,
                           kappa_over_D=self.kappa_over_D if kappa_over_D is None else kappa_over_D,
                           centers=nat.fptr(c), radii=nat.fptr(r))
        for i in range(3):
            g.axes[i] = self.axes[i]
        g.aux = self.aux
        if self.myelin is not None:
            if n_walkers is None:
                raise ValueError("the myelinated cylinder's descriptor needs n_walkers")
            m = self.myelin
            md = nat.Myelin_t()
            for i in range(3):
                md.step_len[i] = m["step_len"][i]
                md.dt_over_T2[i] = m["dt_over_T2"][i]
                md.spin_weight[i] = m["spin_weight"][i]
            for i in range(4):
                md.two_kappa_over_D[i] = m["two_kappa_over_D"][i]
            n0, n1, _ = m["start_counts"](int(n_walkers))
            md.n_start[0] = n0
            md.n_start[1] = n0 + n1
            self.extra["_myelin_keep"] = md
            g.myelin = C.pointer(md)
        if self.packed_myelin is not None:
            if n_walkers is None:
                raise ValueError("the packed myelinated descriptor needs n_walkers")
            m = self.packed_myelin
            n_in, n_my, _ = m["start_counts"](int(n_walkers))
            start = np.ascontiguousarray(
                np.concatenate([[0], np.cumsum(np.concatenate([n_in, n_my]))]), dtype=np.int64)
            pm = nat.PackedMyelin_t(inner_radii=nat.fptr(m["inner"]), step_len=nat.fptr(m["step_len"]),
                                    two_kappa_over_D=nat.fptr(m["two_kappa_over_D"]),
                                    rho_over_D=m["rho_over_D"], myelin_weight=m["myelin_weight"],
                                    start=start.ctypes.data_as(C.POINTER(C.c_int64)))
            for i in range(3):
                pm.dt_over_T2[i] = m["dt_over_T2"][i]
            self.extra["_pmyelin_keep"] = (pm, start)
            g.packed_myelin = C.pointer(pm)
        return g


def _rotation_of(geometry):
    R = getattr(geometry, '_R', None)
    if R is None or getattr(geometry, '_is_identity_rotation', True):
        return None
    return np.asarray(R, dtype=np.float32)


def _myelin_plan(geometry, dt, T1):
    """Constants of make_myelin_step_fn (physics.py:420-512): one step per sample;
    float32 step lengths sqrt(6 D_c dt); 2 kappa / max(D_leaving, 1e-30) per
    membrane side; dt/T2 per compartment (1e6 s for an unset T2 once any is set);
    EPS/NUDGE from the inner radius; T1 on the stored intervals.  The simulate()
    T2 argument does not apply (core.py:485-489)."""
    g = geometry
    dt32 = f32(dt)
    D = [f32(g.D_intra), f32(g.D_myelin), f32(g.D_extra)]
    step = [float(np.sqrt((f32(6.0) * d) * dt32, dtype=np.float32)) for d in D]
    k_in = f32(g.kappa_inner) if g.kappa_inner is not None else f32(0.0)
    k_out = f32(g.kappa_outer) if g.kappa_outer is not None else f32(0.0)
    floor = f32(1e-30)
    two_k = [float(f32(2.0) * (k / np.maximum(D[c], floor)))
             for k, c in ((k_in, 0), (k_in, 1), (k_out, 1), (k_out, 2))]
    t2s = (g.T2_intra, g.T2_myelin, g.T2_extra)
    has_t2 = any(t is not None for t in t2s)
    q2 = [float(dt32 / f32(t if t is not None else 1e6)) for t in t2s]
    has_t1 = T1 is not None
    myelin = dict(step_len=step, two_kappa_over_D=two_k, dt_over_T2=q2,
                  spin_weight=[float(f32(w)) for w in g.water_fractions],
                  start_counts=g.start_counts)
    return WalkPlan(kind=geo.KIND_MYELIN, wall=WALL_REFLECT, n_sub=1,
                    size=float(f32(g.inner_radius)), cell=0.0,
                    eps=float(f32(1e-7 * g.inner_radius)), nudge=float(f32(1e-4 * g.inner_radius)),
                    rho_over_D=0.0, kappa_over_D=0.0, step_len=0.0,
                    gamma_dt_sub=float(f32(GAMMA * dt)), dt_sub=float(dt32),
                    inv_T2=0.0, inv_T1=float(f32(1.0 / T1)) if has_t1 else 0.0,
                    has_t2=has_t2, has_t1=has_t1, has_weight=True,
                    rot=_rotation_of(geometry), aux=float(f32(g.outer_radius)), myelin=myelin)


def _packed_myelin_plan(geometry, dt, T1):
    """Constants of make_packed_myelin_step_fn (physics.py:979-1031) and its walk
    (682-724): wall relaxivity rho = max over the inner/outer walls, rho/D_ref with
    D_ref the largest intra/extra diffusivity and the surface sub-step rule on the
    extra-axonal pore when rho > 0 (else one step per sample); per-axon float32 step
    lengths sqrt(6 D dt_sub) and 2 kappa / max(D_leaving, 1e-30); dt_sub/T2 per class
    from the first axon; NUDGE of the smallest inner radius."""
    g = geometry
    n = g.N_actual
    rho = max(float(np.max(g._rho_inner_np)), float(np.max(g._rho_outer_np)))
    if rho > 0.0:
        D_ref = float(max(np.max(g._D_intra_np), np.max(g._D_extra_np)))
        rho_over_D = float(f32(rho / D_ref))
        n_sub = surface_sub_steps(g, D_ref, dt)
    else:
        rho_over_D, n_sub = 0.0, 1
    dt_sub = dt / n_sub
    dt32 = f32(dt_sub)
    Ds = [g._D_intra_np[:n].astype(np.float32), g._D_myelin_np[:n].astype(np.float32),
          g._D_extra_np[:n].astype(np.float32)]
    step = np.stack([np.sqrt((f32(6.0) * d) * dt32).astype(np.float32) for d in Ds], axis=1)
    floor = f32(1e-30)
    ki = g._kappa_inner_np[:n].astype(np.float32)
    ko = g._kappa_outer_np[:n].astype(np.float32)
    k2 = np.stack([f32(2.0) * (ki / np.maximum(Ds[0], floor)),
                   f32(2.0) * (ki / np.maximum(Ds[1], floor)),
                   f32(2.0) * (ko / np.maximum(Ds[1], floor)),
                   f32(2.0) * (ko / np.maximum(Ds[2], floor))], axis=1).astype(np.float32)
    q2 = [0.0, 0.0, 0.0]
    if g._has_t2:
        q2 = [float(dt32 / f32(t[0])) for t in g._T2_np]
    has_t1 = T1 is not None
    pm = dict(inner=np.ascontiguousarray(g._inner_radii_np[:n], np.float32),
              step_len=np.ascontiguousarray(step), two_kappa_over_D=np.ascontiguousarray(k2),
              dt_over_T2=q2, rho_over_D=rho_over_D,
              myelin_weight=float(f32(g._myelin_proton_density)), start_counts=g.start_counts)
    return WalkPlan(kind=geo.KIND_PACKED_MYELIN, wall=WALL_REFLECT, n_sub=int(n_sub), size=0.0,
                    cell=float(f32(g._L_float)), eps=float(g._eps), nudge=float(g._nudge),
                    rho_over_D=0.0, kappa_over_D=0.0, step_len=0.0,
                    gamma_dt_sub=float(f32(GAMMA * dt_sub)), dt_sub=float(dt32),
                    inv_T2=0.0, inv_T1=float(f32(1.0 / T1)) if has_t1 else 0.0,
                    has_t2=bool(g._has_t2), has_t1=has_t1, has_weight=True,
                    centers=np.ascontiguousarray(g._centers_np[:n], np.float32),
                    radii=np.ascontiguousarray(g._outer_radii_np[:n], np.float32),
                    rot=_rotation_of(geometry), packed_myelin=pm)


def make_walk_plan(geometry, diffusivity, dt, T2=None, T1=None, substep_rule="fused"):
    """Resolve wall routine, sub-steps and float32 constants for one walk.

    ``substep_rule="fused"`` follows ``make_step_fn`` (R/25 permeable, R_char/8
    with surface relaxivity, collision rule otherwise); ``"replay"`` follows the
    walk-once producer (R/6, or R/25 when permeable).
    """
    kind = getattr(geometry, '_walk_kind', None)
    if kind is None:
        raise NotImplementedError(
            f"geometry {type(geometry).__name__} is not supported by the HIP walk")
    if kind == geo.KIND_MYELIN:
        return _myelin_plan(geometry, dt, T1)
    if kind == geo.KIND_PACKED_MYELIN:
        return _packed_myelin_plan(geometry, dt, T1)
    if diffusivity is None:
        raise ValueError("diffusivity is required for standard geometries")
    D = float(diffusivity)
    dt = float(dt)
    has_surf = getattr(geometry, 'surface_relaxivity_t2', None) is not None
    has_perm = getattr(geometry, 'permeability', None) is not None
    has_t2 = T2 is not None
    has_t1 = T1 is not None

    rho_over_D = 0.0
    kappa_over_D = 0.0
    if has_perm:
        wall = WALL_PERMEATE
        kappa_over_D = float(f32(geometry.permeability / D))
        if has_surf:
            rho_over_D = float(f32(geometry.surface_relaxivity_t2 / D))
        n_sub = permeable_sub_steps(geometry, D, dt)
    elif has_surf:
        wall = WALL_SURFACE
        rho_over_D = float(f32(geometry.surface_relaxivity_t2) / f32(D))
        n_sub = surface_sub_steps(geometry, D, dt)
    else:
        wall = WALL_REFLECT
        n_sub = collision_sub_steps(geometry, D, dt)
    if substep_rule == "replay":
        n_sub = replay_sub_steps(geometry, D, dt, permeable=has_perm)
    elif substep_rule != "fused":
        raise ValueError(f"unknown substep_rule {substep_rule!r}")
    if kind == geo.KIND_FREE:
        wall = WALL_REFLECT
    if kind == geo.KIND_BOX1D and wall == WALL_PERMEATE:
        raise NotImplementedError("Box1D has no permeable wall")

    dt_sub = f32(dt / n_sub)
    step_len = float(np.sqrt(f32(6.0 * D) * dt_sub, dtype=np.float32))
    gamma_dt_sub = float(f32(GAMMA * (dt / n_sub)))
    inv_T2 = float(f32(1.0 / T2)) if has_t2 else 0.0
    inv_T1 = float(f32(1.0 / T1)) if has_t1 else 0.0

    size, cell = 0.0, 0.0
    axes, aux = (0.0, 0.0, 0.0), 0.0
    centers = radii = None
    if kind == geo.KIND_BOX1D:
        size = geometry.length
        eps = nudge = 0.0
    elif kind == geo.KIND_ELLIPSOID:
        # geometries.py:1208-1213, 1238-1240: inv_semi_sq = 1/(semi*semi) in float32,
        # EPS/NUDGE from the smallest semi-axis, clamp factor 1 - NUDGE/min_semi.
        semi = np.asarray(geometry.semiaxes, dtype=np.float64)
        s32 = semi.astype(np.float32)
        inv = f32(1.0) / (s32 * s32)
        mn = float(np.min(semi))
        eps = float(f32(1e-7 * mn))
        nudge = float(f32(1e-4 * mn))
        axes = tuple(float(v) for v in inv)
        aux = float(f32(1.0) - f32(nudge) / f32(mn))
    elif kind == geo.KIND_SLAB:
        size = geometry.length
        eps = nudge = 0.0
    elif kind in (geo.KIND_SHELL, geo.KIND_SHELL_CYL):
        size = geometry.r_inner
        aux = float(f32(geometry.r_outer))
        eps = float(f32(1e-7 * geometry.r_inner))
        nudge = float(f32(1e-4 * geometry.r_inner))
    elif kind in (geo.KIND_CYLINDER, geo.KIND_SPHERE):
        size = geometry.radius
        eps = float(f32(1e-7 * geometry.radius))
        nudge = float(f32(1e-4 * geometry.radius))
    elif kind in (geo.KIND_PACKED_CYL, geo.KIND_PACKED_SPH):
        min_r = float(np.min(geometry._radii_np))
        eps = float(f32(1e-7 * min_r))
        nudge = float(f32(1e-4 * min_r))
        cell = geometry._L_float
        centers = np.ascontiguousarray(geometry._centers_np, dtype=np.float32)
        radii = np.ascontiguousarray(geometry._radii_np, dtype=np.float32)
    else:
        eps = nudge = 0.0

    return WalkPlan(kind=kind, wall=wall, n_sub=int(n_sub), size=float(f32(size)),
                    cell=float(f32(cell)), eps=eps, nudge=nudge,
                    rho_over_D=rho_over_D, kappa_over_D=kappa_over_D,
                    step_len=step_len, gamma_dt_sub=gamma_dt_sub, dt_sub=float(dt_sub),
                    inv_T2=inv_T2, inv_T1=inv_T1, has_t2=has_t2, has_t1=has_t1,
                    has_weight=bool(has_surf or has_perm or has_t2 or has_t1),
                    centers=centers, radii=radii, rot=_rotation_of(geometry),
                    axes=axes, aux=aux)
