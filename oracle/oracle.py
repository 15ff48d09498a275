"""Python side of the CPU oracle (test infrastructure only).

The oracle checks the HIP walk; it is imported only by tests/, by
__graft_entry__.smoke() and by bench.py's cpu_baseline leg.

``oracle_plan`` restates, independently of the dmipy_sim_amd package, how the
reference turns (geometry, D, dt, T2, T1) into the step constants:
  sub-step rules     dmipy_sim/physics.py:15-48 (R/25 permeable), 51-69 + 163-188
                     (R_char/8 surface), 72-95 (collision, 1 here)
  step length        physics.py:256-260   sqrt(f32(6 D) * f32(dt/n_sub))
  gamma*dt_sub       physics.py:288-290   f32(GAMMA*dt/n_sub)
  rho/D, kappa/D     physics.py:279-281 (permeable: f32(x/D)), 322-328
                     (surface: f32(rho)/f32(D))
  EPS / NUDGE        geometries.py:503-504, 1694-1696
``run_oracle`` calls walk_oracle.c, which restates the step itself.
"""
import ctypes as C
import math
import os
import subprocess

import numpy as np

GAMMA = 267.513e6
HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle_walk.so")

KIND = {"FreeDiffusion": 0, "Box1D": 1, "Cylinder": 2, "Sphere": 3,
        "PackedCylinders": 4, "PackedSpheres": 5, "Ellipsoid": 6, "PermeableSlab1D": 7,
        "PermeableShell": 8}
W_REFLECT, W_SURFACE, W_PERMEATE = 0, 1, 2

_fp = C.POINTER(C.c_float)
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_up = C.POINTER(C.c_uint32)


class OracleArgs(C.Structure):
    _fields_ = [("kind", C.c_int32), ("wall", C.c_int32), ("n_obj", C.c_int32),
                ("size", C.c_float), ("cell", C.c_float), ("eps", C.c_float),
                ("nudge", C.c_float), ("rho_over_D", C.c_float), ("kappa_over_D", C.c_float),
                ("centers", _fp), ("radii", _fp), ("n_sub", C.c_int32),
                ("step_len", C.c_float), ("gamma_dt_sub", C.c_float), ("dt_sub", C.c_float),
                ("inv_T2", C.c_float), ("inv_T1", C.c_float), ("has_t2", C.c_int32),
                ("has_t1", C.c_int32), ("n_t", C.c_int32), ("n_meas", C.c_int32),
                ("G", _fp), ("chi", _fp), ("n_echo", C.c_int32), ("echo_idx", _ip),
                ("n_walkers", C.c_int64), ("seed", C.c_uint64), ("r0", _fp),
                ("inj_n", _fp), ("inj_u", _fp), ("cell_form", C.c_int32),
                ("phase64", C.c_int32), ("threads", C.c_int32),
                ("walker_begin", C.c_int64), ("walker_end", C.c_int64),
                ("instance", C.c_uint32), ("axes", C.c_float * 3), ("aux", C.c_float),
                ("m_step", C.c_float * 3), ("m_q2", C.c_float * 3), ("m_wf", C.c_float * 3),
                ("m_k2", C.c_float * 4), ("m_n0", C.c_int64), ("m_n01", C.c_int64),
                ("pm_inner", _fp), ("pm_step", _fp), ("pm_k2", _fp), ("pm_q2", C.c_float * 3),
                ("pm_rhoD", C.c_float), ("pm_wmy", C.c_float), ("pm_start", C.POINTER(C.c_int64)),
                ("first_walker", C.c_int64)]


class OracleOut(C.Structure):
    _fields_ = [("signal", _dp), ("signal_im", _dp), ("final_pos", _fp),
                ("comp_origin", _ip), ("comp_final", _ip), ("comp_full", _ip),
                ("n_cross", _up), ("log_w", _fp), ("phi", _fp), ("n_hits", C.c_uint64),
                ("pos_full", _fp), ("dlog_full", _fp), ("occ_full", _fp),
                ("signal_x", _dp), ("signal_x_im", _dp), ("phi_x", _dp)]


_lib = None


def build():
    """Compile walk_oracle.c (make -B in oracle/)."""
    subprocess.run(["make", "-B", "-C", HERE], check=True, stdout=subprocess.DEVNULL)


def _bind(path):
    h = C.CDLL(path)
    h.oracle_walk.restype = C.c_int
    h.oracle_walk.argtypes = [C.POINTER(OracleArgs), C.POINTER(OracleOut)]
    h.oracle_philox4x32_10.restype = None
    h.oracle_philox4x32_10.argtypes = [_up, _up, _up]
    return h


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = _bind(LIB)
    return _lib


def native_lib(out_dir):
    """The same restatement built for speed on the host it runs on (a CPU timing
    baseline, not the checker): -O3 -march=native with FMA contraction allowed,
    OpenMP over walkers.  Compiled into out_dir; returns a bound library."""
    path = os.path.join(out_dir, "liboracle_walk_native.so")
    subprocess.run(["gcc", "-O3", "-march=native", "-fPIC", "-std=c11", "-fopenmp", "-shared",
                    "-o", path, os.path.join(HERE, "walk_oracle.c"), "-lm"], check=True)
    return _bind(path)


def philox4x32_10(ctr, key):
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    o = np.zeros(4, np.uint32)
    lib().oracle_philox4x32_10(c.ctypes.data_as(_up), k.ctypes.data_as(_up), o.ctypes.data_as(_up))
    return o


f32 = np.float32


def _radius(g):
    """Smallest geometric radius (physics.py:15-32)."""
    for attr in ("radius", "sphere_radius"):
        v = getattr(g, attr, None)
        if v is not None:
            return float(v)
    radii = getattr(g, "_radii_np", None)
    if radii is not None and len(radii) > 0:
        return float(np.min(radii))
    return None


def _myelin_plan(g, dt, T1):
    """MyelinatedCylinder (make_myelin_step_fn, physics.py:455-512): one step per
    sample, sqrt(float32(6) D_c dt) steps, 2 kappa / max(D_leaving, 1e-30),
    dt/T2_c with 1e6 s for unset T2s, EPS/NUDGE of R_in; start allotment by
    cross-section area with extra-axonal space bounded at 2 R_out
    (geometries.py:921-946)."""
    dt32 = f32(dt)
    Ds = [f32(g.D_intra), f32(g.D_myelin), f32(g.D_extra)]
    ki = f32(0.0) if g.kappa_inner is None else f32(g.kappa_inner)
    ko = f32(0.0) if g.kappa_outer is None else f32(g.kappa_outer)
    k2 = []
    for k, Dl in ((ki, Ds[0]), (ki, Ds[1]), (ko, Ds[1]), (ko, Ds[2])):
        k2.append(float(f32(2.0) * f32(k / max(Dl, f32(1e-30)))))
    T2s = [g.T2_intra, g.T2_myelin, g.T2_extra]
    Ri, Ro = g.inner_radius, g.outer_radius
    area = [math.pi * Ri ** 2, math.pi * (Ro ** 2 - Ri ** 2), math.pi * ((2 * Ro) ** 2 - Ro ** 2)]
    return dict(kind=10, wall=W_REFLECT, n_sub=1, rho_over_D=0.0, kappa_over_D=0.0, step_len=0.0,
                gamma_dt_sub=float(f32(GAMMA * dt)), dt_sub=float(dt32), inv_T2=0.0,
                inv_T1=float(f32(1.0 / T1)) if T1 is not None else 0.0,
                has_t2=any(t is not None for t in T2s), has_t1=T1 is not None,
                size=float(f32(Ri)), cell=0.0, eps=float(f32(1e-7 * Ri)),
                nudge=float(f32(1e-4 * Ri)), centers=None, radii=None, axes=[0.0, 0.0, 0.0],
                aux=float(f32(Ro)),
                m_step=[float(np.sqrt(f32(f32(6.0) * d) * dt32)) for d in Ds],
                m_q2=[float(f32(dt32 / f32(1e6 if t is None else t))) for t in T2s],
                m_wf=[float(f32(w)) for w in g.water_fractions], m_k2=k2, m_area=area)


def _packed_myelin_plan(g, dt, T1):
    """PackedMyelinatedCylinders (make_packed_myelin_step_fn, physics.py:979-1031, and
    the walk's constants, 708-724): rho = max wall relaxivity, rho/D_ref with D_ref =
    max(D_intra, D_extra) and the surface sub-step rule on the extra-axonal pore
    area/perimeter (physics.py:51-69, 163-188); per-axon sqrt(f32(6) D dt_sub) and
    2 kappa/max(D_leaving, 1e-30); dt_sub/T2 of the first axon per class; NUDGE of
    the smallest inner radius (geometries.py:3067-3070)."""
    n = g.N_actual
    rho = max(float(np.max(g._rho_inner_np)), float(np.max(g._rho_outer_np)))
    n_sub, rD = 1, 0.0
    if rho > 0.0:
        D_ref = float(max(np.max(g._D_intra_np), np.max(g._D_extra_np)))
        rD = float(f32(rho / D_ref))
        frac = getattr(g, "surface_substep_frac", None)
        frac = 8.0 if frac is None else frac
        if frac and frac > 0:
            ro = g._outer_radii_np[g._outer_radii_np > 0]
            area = g._cell_size ** 2 - float(np.sum(math.pi * ro ** 2))
            perim = float(np.sum(2.0 * math.pi * ro))
            Rc = area / perim
            n_sub = max(1, math.ceil(dt / ((Rc / frac) ** 2 / (6.0 * D_ref))))
    dt_sub = dt / n_sub
    d32 = f32(dt_sub)
    step = np.zeros((n, 3), np.float32)
    k2 = np.zeros((n, 4), np.float32)
    for k in range(n):
        Dk = [f32(g._D_intra_np[k]), f32(g._D_myelin_np[k]), f32(g._D_extra_np[k])]
        for c in range(3):
            step[k, c] = np.sqrt(f32(f32(6.0) * Dk[c]) * d32)
        ki, ko = f32(g._kappa_inner_np[k]), f32(g._kappa_outer_np[k])
        for j, (kap, Dl) in enumerate(((ki, Dk[0]), (ki, Dk[1]), (ko, Dk[1]), (ko, Dk[2]))):
            k2[k, j] = f32(2.0) * f32(kap / max(Dl, f32(1e-30)))
    q2 = [0.0, 0.0, 0.0]
    if g._has_t2:
        q2 = [float(f32(d32 / f32(t[0]))) for t in g._T2_np]
    ri = g._inner_radii_np[:n]
    return dict(kind=11, wall=W_REFLECT, n_sub=int(n_sub), rho_over_D=0.0, kappa_over_D=0.0,
                step_len=0.0, gamma_dt_sub=float(f32(GAMMA * dt_sub)), dt_sub=float(d32),
                inv_T2=0.0, inv_T1=float(f32(1.0 / T1)) if T1 is not None else 0.0,
                has_t2=bool(g._has_t2), has_t1=T1 is not None, size=0.0,
                cell=float(f32(g._cell_size)), eps=float(f32(1e-7 * float(ri.min()))),
                nudge=float(f32(1e-4 * float(ri.min()))),
                centers=np.ascontiguousarray(g._centers_np[:n], np.float32),
                radii=np.ascontiguousarray(g._outer_radii_np[:n], np.float32),
                axes=[0.0, 0.0, 0.0], aux=0.0, pm_inner=np.ascontiguousarray(ri, np.float32),
                pm_step=step, pm_k2=k2, pm_q2=q2, pm_rhoD=rD,
                pm_wmy=float(f32(g._myelin_proton_density)), pm_geom=g)


def packed_myelin_start(p, N):
    """Start boundaries: intra of each axon, myelin of each axon (area allotment,
    geometries.py:3112-3137), the remainder extra-axonal."""
    g = p["pm_geom"]
    n = g.N_actual
    L = g._cell_size
    ri, ro = g._inner_radii_np[:n], g._outer_radii_np[:n]
    a_in = [math.pi * r ** 2 for r in ri]
    a_my = [math.pi * (o ** 2 - r ** 2) for r, o in zip(ri, ro)]
    tot = (L * L - math.pi * float(np.sum(ro ** 2))) + sum(a_in) + sum(a_my)
    n_in = [int(round(N * a / tot)) for a in a_in]
    n_my = [int(round(N * a / tot)) for a in a_my]
    excess = sum(n_in) + sum(n_my) - N
    k = 0
    while excess > 0 and k < n:
        t = min(excess, n_in[k])
        n_in[k] -= t
        excess -= t
        k += 1
    return np.ascontiguousarray(np.cumsum([0] + n_in + n_my), dtype=np.int64)


def myelin_start_counts(p, N):
    """Walkers starting intra-axonal and in the sheath (the rest start outside)."""
    a = p["m_area"]
    tot = a[0] + a[1] + a[2]
    return int(round(N * a[0] / tot)), int(round(N * a[1] / tot))


def oracle_plan(geom, D, dt, T2=None, T1=None):
    """Step constants of one walk, restated from the reference (see module doc)."""
    name = type(geom).__name__
    if name == "MyelinatedCylinder":
        return _myelin_plan(geom, dt, T1)
    if name == "PackedMyelinatedCylinders":
        return _packed_myelin_plan(geom, dt, T1)
    kind = KIND[name]
    if name == "PermeableShell" and geom.kind == "cylinder":
        kind = 9
    has_perm = getattr(geom, "permeability", None) is not None
    has_surf = getattr(geom, "surface_relaxivity_t2", None) is not None
    R = _radius(geom)
    if has_perm:
        wall = W_PERMEATE
        n_sub = 1 if R is None else max(1, math.ceil(dt / (R * R / (3750.0 * D))))
        kD = float(f32(geom.permeability / D))
        rD = float(f32(geom.surface_relaxivity_t2 / D)) if has_surf else 0.0
    elif has_surf:
        wall = W_SURFACE
        frac = getattr(geom, "surface_substep_frac", None)
        frac = 8.0 if frac is None else frac
        if not frac or frac <= 0 or R is None:
            n_sub = 1
        else:
            n_sub = max(1, math.ceil(dt / ((R / frac) ** 2 / (6.0 * D))))
        kD = 0.0
        rD = float(f32(f32(geom.surface_relaxivity_t2) / f32(D)))
    else:
        wall, n_sub, kD, rD = W_REFLECT, 1, 0.0, 0.0
    if kind == 0:
        wall = W_REFLECT
    dt_sub = f32(dt / n_sub)
    step = float(np.sqrt(f32(f32(6.0 * D) * dt_sub)))
    p = dict(kind=kind, wall=wall, n_sub=int(n_sub), rho_over_D=rD, kappa_over_D=kD,
             step_len=step, gamma_dt_sub=float(f32(GAMMA * (dt / n_sub))), dt_sub=float(dt_sub),
             inv_T2=float(f32(1.0 / T2)) if T2 is not None else 0.0,
             inv_T1=float(f32(1.0 / T1)) if T1 is not None else 0.0,
             has_t2=T2 is not None, has_t1=T1 is not None, size=0.0, cell=0.0,
             eps=0.0, nudge=0.0, centers=None, radii=None, axes=[0.0, 0.0, 0.0], aux=0.0)
    if kind == 1:
        p["size"] = float(f32(geom.length))
    elif kind in (2, 3):
        p["size"] = float(f32(geom.radius))
        p["eps"] = float(f32(1e-7 * geom.radius))
        p["nudge"] = float(f32(1e-4 * geom.radius))
    elif kind == 6:
        # Ellipsoid (geometries.py:1208-1213, 1238-1240)
        semi = np.asarray(geom.semiaxes, np.float64)
        s32 = semi.astype(np.float32)
        inv = np.float32(1.0) / (s32 * s32)
        mn = float(semi.min())
        p["eps"] = float(f32(1e-7 * mn))
        p["nudge"] = float(f32(1e-4 * mn))
        p["axes"] = [float(v) for v in inv]
        p["aux"] = float(np.float32(1.0) - np.float32(p["nudge"]) / np.float32(mn))
    elif kind == 7:
        p["size"] = float(f32(geom.length))
    elif kind in (8, 9):
        p["size"] = float(f32(geom.r_inner))
        p["aux"] = float(f32(geom.r_outer))
        p["eps"] = float(f32(1e-7 * geom.r_inner))
        p["nudge"] = float(f32(1e-4 * geom.r_inner))
    elif kind in (4, 5):
        rmin = float(np.min(geom._radii_np))
        p["eps"] = float(f32(1e-7 * rmin))
        p["nudge"] = float(f32(1e-4 * rmin))
        p["cell"] = float(f32(geom._L_float))
        p["centers"] = np.ascontiguousarray(geom._centers_np, dtype=np.float32)
        p["radii"] = np.ascontiguousarray(geom._radii_np, dtype=np.float32)
    return p


def run_oracle(geom, D, waveform, n_walkers, seed=0, T2=None, T1=None, r0=None,
               inject_normals=None, inject_uniforms=None, cell_form=1, phase64=False,
               threads=0, echo_indices=None, walker_range=None, plan=None, G_native=None,
               instance=0, per_sample=False, first_walker=0, exact_phase=False, library=None):
    """Walk on the CPU; returns a dict like dmipy_sim_amd.walk (native frame).

    phase64: False = float32 phase (the reference's arithmetic), True = float32
    increments summed in float64.  exact_phase: also return 'signal_x',
    'signal_x_im' and 'phi_x', the same walk with every phase increment formed and
    summed in float64 from the float32 positions (the exact phase of the float32
    trajectory: the FP64 oracle the kernel's phase error is stated against).
    first_walker: global index of walker 0 (a shard of a larger walk).  library: a
    bound library from native_lib (default: the checker build)."""
    p = plan if plan is not None else oracle_plan(geom, D, waveform.dt, T2, T1)
    G = np.ascontiguousarray(waveform.G if G_native is None else G_native, dtype=np.float32)
    n_meas, n_t, _ = G.shape
    N = int(n_walkers)
    chi = getattr(waveform, "chi_perp", None)
    chi = None if chi is None else np.ascontiguousarray(chi, dtype=np.float32).reshape(n_t)
    echo = None if echo_indices is None else np.ascontiguousarray(echo_indices, dtype=np.int32)
    r0 = None if r0 is None else np.ascontiguousarray(r0, dtype=np.float32).reshape(N, 3)
    inj_n = None if inject_normals is None else np.ascontiguousarray(inject_normals, dtype=np.float32)
    inj_u = None if inject_uniforms is None else np.ascontiguousarray(inject_uniforms, dtype=np.float32)
    fp = lambda a: None if a is None else a.ctypes.data_as(_fp)
    w0, w1 = (0, 0) if walker_range is None else walker_range
    A = OracleArgs(kind=p["kind"], wall=p["wall"],
                   n_obj=0 if p["radii"] is None else len(p["radii"]),
                   size=p["size"], cell=p["cell"], eps=p["eps"], nudge=p["nudge"],
                   rho_over_D=p["rho_over_D"], kappa_over_D=p["kappa_over_D"],
                   centers=fp(p["centers"]), radii=fp(p["radii"]), n_sub=p["n_sub"],
                   step_len=p["step_len"], gamma_dt_sub=p["gamma_dt_sub"], dt_sub=p["dt_sub"],
                   inv_T2=p["inv_T2"], inv_T1=p["inv_T1"], has_t2=int(p["has_t2"]),
                   has_t1=int(p["has_t1"]), n_t=n_t, n_meas=n_meas, G=fp(G), chi=fp(chi),
                   n_echo=0 if echo is None else len(echo),
                   echo_idx=None if echo is None else echo.ctypes.data_as(_ip),
                   n_walkers=N, seed=int(seed) & (2**64 - 1), r0=fp(r0), inj_n=fp(inj_n),
                   inj_u=fp(inj_u), cell_form=int(cell_form), phase64=int(bool(phase64)),
                   threads=int(threads), walker_begin=int(w0), walker_end=int(w1),
                   instance=int(instance), aux=float(p.get("aux", 0.0)),
                   first_walker=int(first_walker))
    for k in range(3):
        A.axes[k] = float(p.get("axes", [0.0, 0.0, 0.0])[k])
    if p["kind"] == 10:
        for k in range(3):
            A.m_step[k] = p["m_step"][k]
            A.m_q2[k] = p["m_q2"][k]
            A.m_wf[k] = p["m_wf"][k]
        for k in range(4):
            A.m_k2[k] = p["m_k2"][k]
        n0, n1 = myelin_start_counts(p, N)
        A.m_n0, A.m_n01 = n0, n0 + n1
    keep = []
    if p["kind"] == 11:
        start = packed_myelin_start(p, N)
        keep += [start, p["pm_step"], p["pm_k2"], p["pm_inner"]]
        A.pm_inner = fp(p["pm_inner"])
        A.pm_step = fp(np.ascontiguousarray(p["pm_step"]))
        A.pm_k2 = fp(np.ascontiguousarray(p["pm_k2"]))
        for k in range(3):
            A.pm_q2[k] = p["pm_q2"][k]
        A.pm_rhoD = p["pm_rhoD"]
        A.pm_wmy = p["pm_wmy"]
        A.pm_start = start.ctypes.data_as(C.POINTER(C.c_int64))
    rows = 1 if echo is None else len(echo)
    res = {"signal": np.zeros(rows * n_meas), "signal_im": np.zeros(rows * n_meas),
           "final_pos": np.zeros((N, 3), np.float32), "comp_origin": np.zeros(N, np.int32),
           "comp_final": np.zeros(N, np.int32), "n_cross": np.zeros(N, np.uint32),
           "log_w": np.zeros(N, np.float32), "phi": np.zeros((N, n_meas), np.float32)}
    O = OracleOut(signal=res["signal"].ctypes.data_as(_dp),
                  signal_im=res["signal_im"].ctypes.data_as(_dp),
                  final_pos=fp(res["final_pos"]),
                  comp_origin=res["comp_origin"].ctypes.data_as(_ip),
                  comp_final=res["comp_final"].ctypes.data_as(_ip), comp_full=None,
                  n_cross=res["n_cross"].ctypes.data_as(_up), log_w=fp(res["log_w"]),
                  phi=fp(res["phi"]))
    if per_sample:
        res["pos_full"] = np.zeros((N, n_t, 3), np.float32)
        res["dlog_full"] = np.zeros((N, n_t), np.float32)
        res["occ_full"] = np.zeros((N, n_t), np.float32)
        O.pos_full = fp(res["pos_full"])
        O.dlog_full = fp(res["dlog_full"])
        O.occ_full = fp(res["occ_full"])
        res["comp_full"] = np.zeros((N, n_t), np.int32)
        O.comp_full = res["comp_full"].ctypes.data_as(_ip)
    if exact_phase:
        res["signal_x"] = np.zeros(rows * n_meas)
        res["signal_x_im"] = np.zeros(rows * n_meas)
        res["phi_x"] = np.zeros((N, n_meas))
        O.signal_x = res["signal_x"].ctypes.data_as(_dp)
        O.signal_x_im = res["signal_x_im"].ctypes.data_as(_dp)
        O.phi_x = res["phi_x"].ctypes.data_as(_dp)
    rc = (library or lib()).oracle_walk(C.byref(A), C.byref(O))
    if rc != 0:
        raise RuntimeError("oracle_walk failed")
    if echo is not None:
        for k in ("signal", "signal_im", "signal_x", "signal_x_im"):
            if k in res:
                res[k] = res[k].reshape(rows, n_meas)
    res["n_hits"] = int(O.n_hits)
    res["n_sub"] = p["n_sub"]
    return res
