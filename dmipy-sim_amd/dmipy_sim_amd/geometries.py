"""Substrate geometries: descriptors of the walls the walk reflects from.

The classes keep the constructor signatures and attributes of
dmipy_sim/geometries.py (``radius``, ``orientation``, ``surface_relaxivity_t2``,
``permeability``, ``_radii_np``, ``_L_float``, ``min_gap``, ``init_positions``,
``volume``/``surface_area``/``volume_fraction``) so that code written against the
reference builds the same substrate.  The wall physics itself (specular
reflection, Brownstein-Tarr weights, Powles crossing, periodic images) runs only in
the HIP walk kernel; these objects carry the parameters to it.

* ``FreeDiffusion``   -- dmipy_sim/geometries.py:34-45
* ``Box1D``           -- dmipy_sim/geometries.py:48-140
* ``Sphere``          -- dmipy_sim/geometries.py:143-424
* ``Cylinder``        -- dmipy_sim/geometries.py:427-844
* ``pack_cylinders``  -- dmipy_sim/geometries.py:1502-1601
* ``PackedCylinders`` -- dmipy_sim/geometries.py:1604-2242
* ``pack_spheres``    -- dmipy_sim/geometries.py:2249-2340
* ``PackedSpheres``   -- dmipy_sim/geometries.py:2343-2786
* ``Ellipsoid``       -- dmipy_sim/geometries.py:1137-1425
* ``MyelinatedCylinder`` -- dmipy_sim/geometries.py:847-1136
* ``pack_myelinated_cylinders``, ``PackedMyelinatedCylinders`` -- dmipy_sim/geometries.py:2787-3220
* ``PermeableSlab1D`` -- dmipy_sim/geometries.py:3221-3300
* ``PermeableShell``  -- dmipy_sim/geometries.py:3303-3431
"""
import numpy as np

# Geometry kind codes shared with include/dmipy_walk.h (enum dmw_geom_kind).
KIND_FREE, KIND_BOX1D, KIND_CYLINDER, KIND_SPHERE, KIND_PACKED_CYL, KIND_PACKED_SPH = range(6)
KIND_ELLIPSOID, KIND_SLAB, KIND_SHELL, KIND_SHELL_CYL, KIND_MYELIN, KIND_PACKED_MYELIN = range(6, 12)

# Water per unit myelin-sheath volume, the default myelin proton density of
# MyelinatedCylinder.water_fractions (dmipy_sim/substrate/biophysical_constants.py:523-530).
MYELIN_WATER_PROTON_DENSITY = 0.40


def _rng_from_key(key):
    """numpy Generator for host-side placement: an int seed, a Generator, or any
    object convertible to an int (the reference derives its numpy seed from a JAX
    key, dmipy_sim/geometries.py:188)."""
    if isinstance(key, np.random.Generator):
        return key
    if key is None:
        key = 0
    arr = np.asarray(key).ravel()
    seed = [int(v) & 0xFFFFFFFF for v in arr] if arr.size > 1 else int(arr[0]) & (2**63 - 1)
    return np.random.default_rng(seed)


def _rotation_to_z(v):
    """Rotation matrix R with R @ v = [0, 0, 1] (Rodrigues' formula; the parallel
    and anti-parallel cases are handled explicitly, dmipy_sim/geometries.py:1466-1495)."""
    v = np.asarray(v, dtype=np.float64)
    v = v / np.linalg.norm(v)
    k = np.array([0.0, 0.0, 1.0])
    c = float(np.dot(v, k))
    if abs(c - 1.0) < 1e-10:
        return np.eye(3)
    if abs(c + 1.0) < 1e-10:
        return np.diag([1.0, -1.0, -1.0])
    axis = np.cross(v, k)
    axis = axis / np.linalg.norm(axis)
    ang = np.arccos(np.clip(c, -1.0, 1.0))
    K = np.array([[0.0, -axis[2], axis[1]],
                  [axis[2], 0.0, -axis[0]],
                  [-axis[1], axis[0], 0.0]])
    return np.eye(3) + np.sin(ang) * K + (1.0 - np.cos(ang)) * (K @ K)


class Geometry:
    """Base of the walker substrates: ``init_positions(n_walkers, key)`` gives the
    reference start distribution (dmipy_sim/geometries.py:14-31)."""
    kind = None
    surface_relaxivity_t2 = None
    permeability = None

    def init_positions(self, n_walkers, key=0):
        raise NotImplementedError


def _opt_float(x):
    return float(x) if x is not None else None


class FreeDiffusion(Geometry):
    """Unbounded free diffusion; walkers start at the origin, compartment 0."""
    _walk_kind = KIND_FREE

    def init_positions(self, n_walkers, key=0):
        return np.zeros((n_walkers, 3), dtype=np.float32)


class Box1D(Geometry):
    """Reflecting slab between x=0 and x=length, free along y and z.

    ``surface_relaxivity_t2`` (m/s) applies the Brownstein-Tarr weight
    exp(-2 rho d_perp / D) with d_perp the overshoot past the wall.
    """
    _walk_kind = KIND_BOX1D

    def __init__(self, length, surface_relaxivity_t2=None):
        self.length = float(length)
        self.surface_relaxivity_t2 = _opt_float(surface_relaxivity_t2)

    def volume(self):
        return self.length

    def surface_area(self):
        return 2.0

    def init_positions(self, n_walkers, key=0):
        rng = _rng_from_key(key)
        x = rng.uniform(0.0, self.length, (n_walkers, 1))
        return np.concatenate([x, np.zeros((n_walkers, 2))], axis=1).astype(np.float32)


class Sphere(Geometry):
    """Sphere of ``radius`` at the origin, reflecting (optionally relaxing via
    ``surface_relaxivity_t2``) or permeable (``permeability``, Powles crossing)."""
    _walk_kind = KIND_SPHERE

    def __init__(self, radius, surface_relaxivity_t2=None, permeability=None):
        self.radius = float(radius)
        self.surface_relaxivity_t2 = _opt_float(surface_relaxivity_t2)
        self.permeability = _opt_float(permeability)

    def volume(self):
        return (4.0 / 3.0) * np.pi * self.radius ** 3

    def surface_area(self):
        return 4.0 * np.pi * self.radius ** 2

    def init_positions(self, n_walkers, key=0):
        rng = _rng_from_key(key)
        out, have = [], 0
        while have < n_walkers:
            pts = rng.uniform(-self.radius, self.radius, (n_walkers * 4, 3))
            pts = pts[np.linalg.norm(pts, axis=1) < self.radius]
            out.append(pts)
            have += len(pts)
        return np.concatenate(out)[:n_walkers].astype(np.float32)


class Cylinder(Geometry):
    """Infinite cylinder of ``radius`` along ``orientation`` through the origin.

    The walk runs in the cylinder's native frame (axis = z); ``_R`` maps lab
    vectors into it (dmipy_sim/geometries.py:458-468).
    """
    _walk_kind = KIND_CYLINDER

    def __init__(self, radius, orientation, surface_relaxivity_t2=None, permeability=None):
        self.radius = float(radius)
        o = np.asarray(orientation, dtype=np.float64)
        self.orientation = (o / np.linalg.norm(o)).astype(np.float32)
        R = _rotation_to_z(self.orientation)
        self._R = R.astype(np.float32)
        self._R_inv = R.T.astype(np.float32)
        self._is_identity_rotation = bool(np.allclose(R, np.eye(3)))
        self.surface_relaxivity_t2 = _opt_float(surface_relaxivity_t2)
        self.permeability = _opt_float(permeability)

    def volume(self, L=1.0):
        return np.pi * self.radius ** 2 * float(L)

    def surface_area(self, L=1.0, include_caps=False):
        a = 2.0 * np.pi * self.radius * float(L)
        if include_caps:
            a += 2.0 * np.pi * self.radius ** 2
        return a

    def init_positions(self, n_walkers, key=0):
        rng = _rng_from_key(key)
        out, have = [], 0
        while have < n_walkers:
            xy = rng.uniform(-self.radius, self.radius, (n_walkers * 2, 2))
            xy = xy[np.linalg.norm(xy, axis=1) < self.radius]
            out.append(xy)
            have += len(xy)
        xy = np.concatenate(out)[:n_walkers].astype(np.float32)
        r_cyl = np.stack([xy[:, 0], xy[:, 1], np.zeros(n_walkers, np.float32)], axis=1)
        return (self._R_inv.astype(np.float64) @ r_cyl.T).T.astype(np.float32)


def _check_pack_inputs(radii, target_vf, L, dim):
    radii = np.asarray(radii, dtype=np.float64).ravel()
    if len(radii) == 0:
        raise ValueError("radii must contain at least one element.")
    if np.any(radii <= 0):
        raise ValueError("All radii must be positive.")
    if (target_vf is None) == (L is None):
        raise ValueError("Provide exactly one of target_vf or L.")
    if target_vf is not None:
        if not 0.0 < float(target_vf) < 1.0:
            raise ValueError(f"target_vf must be in (0, 1), got {target_vf}.")
        if dim == 2:
            L = float(np.sqrt(np.pi * np.sum(radii ** 2) / float(target_vf)))
        else:
            L = float(((4.0 / 3.0) * np.pi * np.sum(radii ** 3) / float(target_vf)) ** (1.0 / 3.0))
    else:
        L = float(L)
    if np.any(radii > L / 2.0):
        raise ValueError(
            f"Largest radius ({np.max(radii) * 1e6:.2f} µm) exceeds L/2 "
            f"({L / 2.0 * 1e6:.2f} µm).  Reduce target_vf or supply a larger L.")
    return radii, L


def _rsa(radii, L, seed, max_attempts, dim, what):
    """Random sequential addition in a periodic box, largest objects first, with a
    minimum-image overlap test (dmipy_sim/geometries.py:1569-1595, 2309-2334)."""
    rng = np.random.default_rng(int(seed))
    order = np.argsort(radii)[::-1]
    rs = radii[order]
    cs = np.zeros((len(radii), dim))
    for i, r_new in enumerate(rs):
        for _ in range(max_attempts):
            c_new = rng.uniform(-L / 2.0, L / 2.0, dim)
            ok = True
            for j in range(i):
                dq = c_new - cs[j]
                dq -= L * np.round(dq / L)
                if np.linalg.norm(dq) < r_new + rs[j]:
                    ok = False
                    break
            if ok:
                cs[i] = c_new
                break
        else:
            raise RuntimeError(
                f"RSA failed after {max_attempts} attempts placing {what} {i} "
                f"(r = {r_new * 1e6:.2f} µm).  "
                f"The target packing fraction may exceed what RSA can achieve; "
                f"try reducing target_vf or increasing max_attempts.")
    out = np.empty_like(cs)
    out[order] = cs
    return out


def pack_cylinders(radii, target_vf=None, L=None, seed=0, max_attempts=100_000):
    """Pack parallel cylinders in a periodic square by RSA
    (dmipy_sim/geometries.py:1502-1601).  Returns ``(centers (N,2), L, achieved_vf)``."""
    radii, L = _check_pack_inputs(radii, target_vf, L, 2)
    centers = _rsa(radii, L, seed, max_attempts, 2, "cylinder")
    return centers, L, float(np.pi * np.sum(radii ** 2) / L ** 2)


def pack_spheres(radii, target_vf=None, L=None, seed=0, max_attempts=100_000):
    """Pack spheres in a periodic cube by RSA (dmipy_sim/geometries.py:2249-2340).
    Returns ``(centers (N,3), L, achieved_vf)``."""
    radii, L = _check_pack_inputs(radii, target_vf, L, 3)
    centers = _rsa(radii, L, seed, max_attempts, 3, "sphere")
    return centers, L, float((4.0 / 3.0) * np.pi * np.sum(radii ** 3) / L ** 3)


def pack_dense(radii, target_vf, seed=0, dim=2, gap=0.02, max_iter=200_000):
    """Non-overlapping periodic packing at volume fractions beyond RSA's reach.

    Random start in the final box, then repeated pairwise overlap relaxation: every
    pair closer than ``r_i + r_j + gap*min(r)`` (minimum image) is pushed apart
    along its centre line, each by half the deficit, until no pair overlaps.  The
    reference's RSA stops near VF 0.55 for polydisperse disks
    (dmipy_sim/geometries.py:1508-1509); the relaxation reaches the VF 0.6 packing
    the PackedCylinders workload asks for.  Returns ``(centers, L, achieved_vf)``.
    """
    radii, L = _check_pack_inputs(radii, target_vf, None, dim)
    rng = np.random.default_rng(int(seed))
    n = len(radii)
    c = rng.uniform(-L / 2.0, L / 2.0, (n, dim))
    sep = radii[:, None] + radii[None, :] + gap * float(radii.min())
    iu = np.triu_indices(n, 1)
    for it in range(max_iter):
        d = c[:, None, :] - c[None, :, :]
        d -= L * np.round(d / L)
        dist = np.sqrt(np.sum(d * d, axis=2))
        deficit = sep - dist
        np.fill_diagonal(deficit, -1.0)
        if deficit[iu].max() <= 0.0:
            break
        push = np.where(deficit > 0, 0.5 * deficit * 1.0001 + 1e-12, 0.0)
        unit = d / np.maximum(dist, 1e-30)[:, :, None]
        c += np.sum(push[:, :, None] * unit, axis=1)
        c -= L * np.floor(c / L + 0.5)
    else:
        raise RuntimeError("pack_dense: overlap relaxation did not converge; "
                           "reduce target_vf or gap")
    if dim == 2:
        vf = float(np.pi * np.sum(radii ** 2) / L ** 2)
    else:
        vf = float((4.0 / 3.0) * np.pi * np.sum(radii ** 3) / L ** 3)
    return c, L, vf


def pack_myelinated_cylinders(inner_radii, g_ratios, target_packing, cell_size=None, seed=0,
                              max_attempts=100_000):
    """RSA placement of myelinated cylinders (outer radius = inner / g-ratio) in a
    periodic square, outer boundaries disjoint (dmipy_sim/geometries.py:2787-2857).
    ``cell_size`` defaults to sqrt(pi sum R_out^2 / target_packing).  Returns
    ``(inner_radii, g_ratios, centers (N, 2))``."""
    inner = np.asarray(inner_radii, dtype=np.float64).ravel()
    n = len(inner)
    g = np.broadcast_to(np.asarray(g_ratios, dtype=np.float64).ravel(), (n,)).copy()
    outer = inner / g
    if cell_size is None:
        if target_packing is None:
            raise ValueError("Provide either cell_size or target_packing.")
        cell_size = float(np.sqrt(np.pi * np.sum(outer ** 2) / target_packing))
    centers = _rsa(outer, float(cell_size), seed, max_attempts, 2, "cylinder")
    return inner, g, centers


class PackedMyelinatedCylinders:
    """Periodic square cell of myelinated axons sharing one axis
    (dmipy_sim/geometries.py:2860-3220; fused step physics.py:682-1062).

    Compartments: 0 extra-axonal, 1..N_max intra-axonal of axon k (k+1), N_max+1..2
    N_max myelin of axon k; ``simulate`` reports them as 0 intra, 1 myelin, 2 extra.
    Per-axon diffusivities, permeabilities and wall relaxivities broadcast from
    scalars; T2 per compartment class (the first axon's values apply); myelin water
    carries the spin weight ``_myelin_proton_density``.  Walker placement, the
    walk and the surface-relaxivity sub-steps run in the HIP kernel.
    """
    _is_packed_myelinated = True
    _walk_kind = KIND_PACKED_MYELIN

    def __init__(self, inner_radii, g_ratios, centers, cell_size, N_max=128,
                 orientation=(0., 0., 1.), D_intra=2e-9, D_myelin=0.0, D_extra=2e-9,
                 T2_intra=None, T2_myelin=None, T2_extra=None, kappa_inner=0.0,
                 kappa_outer=0.0, rho_inner=0.0, rho_outer=0.0):
        inner = np.asarray(inner_radii, dtype=np.float64).ravel()
        n = len(inner)
        g = np.broadcast_to(np.asarray(g_ratios, dtype=np.float64).ravel(), (n,)).copy()
        centers = np.asarray(centers, dtype=np.float64)
        if centers.shape != (n, 2):
            raise ValueError(f"centers shape {centers.shape} must be ({n}, 2)")
        if N_max < n:
            raise ValueError(f"N_max={N_max} < N_actual={n}")
        self.N_actual = n
        self.N_max = int(N_max)
        self._cell_size = float(cell_size)
        self._L_float = float(cell_size)
        self._myelin_proton_density = MYELIN_WATER_PROTON_DENSITY

        def bcast(x, name, opt=False, zero_none=False):
            if x is None:
                if opt:
                    return None
                x = 0.0
            x = np.asarray(x, dtype=np.float64).ravel()
            if x.size == 1:
                x = np.broadcast_to(x, (n,)).copy()
            elif x.size != n:
                raise ValueError(f"{name} must be scalar or length {n}, got {x.size}")
            out = np.zeros(self.N_max)
            out[:n] = x
            return out

        self._inner_radii_np = bcast(inner, 'inner_radii')
        self._outer_radii_np = bcast(inner / g, 'outer_radii')
        c = np.zeros((self.N_max, 2))
        c[:n] = centers
        self._centers_np = c
        self._D_intra_np = bcast(D_intra, 'D_intra')
        self._D_myelin_np = bcast(D_myelin, 'D_myelin')
        self._D_extra_np = bcast(D_extra, 'D_extra')
        self._kappa_inner_np = bcast(kappa_inner, 'kappa_inner')
        self._kappa_outer_np = bcast(kappa_outer, 'kappa_outer')
        self._rho_inner_np = bcast(rho_inner, 'rho_inner')
        self._rho_outer_np = bcast(rho_outer, 'rho_outer')
        self._has_t2 = T2_intra is not None or T2_myelin is not None or T2_extra is not None
        big = np.full(self.N_max, 1e6)
        self._T2_np = tuple(
            (bcast(t, name, opt=True) if t is not None else big) if self._has_t2 else None
            for t, name in ((T2_intra, 'T2_intra'), (T2_myelin, 'T2_myelin'),
                            (T2_extra, 'T2_extra')))
        o = np.asarray(orientation, dtype=np.float64)
        self.orientation = (o / np.linalg.norm(o)).astype(np.float32)
        R = _rotation_to_z(self.orientation)
        self._R = R.astype(np.float32)
        self._R_inv = R.T.astype(np.float32)
        self._is_identity_rotation = bool(np.allclose(R, np.eye(3)))
        nz = self._inner_radii_np[self._inner_radii_np > 0]
        ref_r = float(np.min(nz)) if len(nz) else 1e-6
        self._eps = np.float32(1e-7 * ref_r)
        self._nudge = np.float32(1e-4 * ref_r)
        self.min_gap = _min_gap(self._centers_np[:n], self._outer_radii_np[:n], self._L_float)

    def volume_fraction(self, compartment):
        L, n = self._L_float, self.N_actual
        ri, ro = self._inner_radii_np[:n], self._outer_radii_np[:n]
        if compartment == 'intra':
            return float(np.pi * np.sum(ri ** 2) / L ** 2)
        if compartment == 'myelin':
            return float(np.pi * np.sum(ro ** 2 - ri ** 2) / L ** 2)
        if compartment == 'extra':
            return float((L ** 2 - np.pi * np.sum(ro ** 2)) / L ** 2)
        raise ValueError(
            f"compartment must be 'intra', 'myelin', or 'extra'; got {compartment!r}")

    def start_counts(self, n_walkers):
        """Walkers per compartment by area, in placement order: intra of each axon,
        myelin of each axon, then extra-axonal (the remainder; a negative remainder
        is trimmed from the intra counts in axon order)."""
        L, n = self._L_float, self.N_actual
        ri, ro = self._inner_radii_np[:n], self._outer_radii_np[:n]
        a_in = np.pi * ri ** 2
        a_my = np.pi * (ro ** 2 - ri ** 2)
        a_ex = L * L - np.pi * np.sum(ro ** 2)
        tot = a_ex + np.sum(a_in) + np.sum(a_my)
        n_in = np.array([int(round(n_walkers * a / tot)) for a in a_in], dtype=np.int64)
        n_my = np.array([int(round(n_walkers * a / tot)) for a in a_my], dtype=np.int64)
        n_ex = n_walkers - int(n_in.sum()) - int(n_my.sum())
        if n_ex < 0:
            excess = -n_ex
            for k in range(n):
                trim = min(excess, int(n_in[k]))
                n_in[k] -= trim
                excess -= trim
                if excess == 0:
                    break
            n_ex = 0
        return n_in, n_my, n_ex

    def init_positions(self, n_walkers, key=0):
        """Host placement in the cell (lab frame, z = 0) with the area allotment;
        sets ``_init_compartments`` (0 extra, k+1 intra, N_max+k+1 myelin).  The GPU
        walk places its walkers itself."""
        rng = _rng_from_key(key)
        n, L = self.N_actual, self._L_float
        ri, ro, c = self._inner_radii_np[:n], self._outer_radii_np[:n], self._centers_np[:n]
        n_in, n_my, _ = self.start_counts(n_walkers)
        pos = np.zeros((n_walkers, 3), np.float32)
        comp = np.zeros(n_walkers, np.int32)
        i = 0
        for counts, lo_r, hi_r, base in ((n_in, None, ri, 1), (n_my, ri, ro, self.N_max + 1)):
            for k in range(n):
                nk = int(counts[k])
                if nk == 0:
                    continue
                got, have = [], 0
                while have < nk:
                    xy = rng.uniform(-hi_r[k], hi_r[k], (max(nk * 4, 64), 2))
                    d = np.linalg.norm(xy, axis=1)
                    keep = d < hi_r[k] if lo_r is None else (d >= lo_r[k]) & (d < hi_r[k])
                    got.append(xy[keep])
                    have += int(keep.sum())
                pos[i:i + nk, :2] = np.concatenate(got)[:nk] + c[k]
                comp[i:i + nk] = base + k
                i += nk
        m = n_walkers - i
        got, have = [], 0
        while have < m:
            xy = rng.uniform(-L / 2.0, L / 2.0, (max(m * 4, 1024), 2))
            out = np.ones(len(xy), bool)
            for k in range(n):
                d = xy - c[k]
                d -= L * np.round(d / L)
                out &= np.sum(d ** 2, axis=1) > ro[k] ** 2
            got.append(xy[out])
            have += int(out.sum())
        if m > 0:
            pos[i:, :2] = np.concatenate(got)[:m]
        self._init_compartments = comp
        return (self._R_inv.astype(np.float64) @ pos.T.astype(np.float64)).T.astype(np.float32)


def _min_gap(centers, radii, L):
    """Smallest clear gap between two object surfaces, periodic images included
    (dmipy_sim/geometries.py:1700-1715)."""
    n = len(radii)
    g = float('inf')
    for i in range(n):
        for j in range(i + 1, n):
            dq = centers[i] - centers[j]
            dq -= L * np.round(dq / L)
            g = min(g, float(np.linalg.norm(dq) - radii[i] - radii[j]))
    for i in range(n):
        g = min(g, L - 2.0 * radii[i])
    return float(g)


def _outside_all(pts, centers, radii, L):
    ok = np.ones(len(pts), dtype=bool)
    for k in range(len(radii)):
        d = pts - centers[k]
        d -= L * np.round(d / L)
        ok &= np.sum(d * d, axis=1) > radii[k] ** 2
    return ok


class PackedCylinders(Geometry):
    """Parallel cylinders in a periodic square cell of side ``L``.

    Walkers start in the extra-axonal space; positions stay unfolded for the phase
    while wall tests use the minimum image.  Compartments: 0 extra-axonal, k inside
    cylinder k (1-based).
    """
    _walk_kind = KIND_PACKED_CYL

    def __init__(self, radii, centers, L, orientation=(0., 0., 1.),
                 surface_relaxivity_t2=None, permeability=None):
        radii = np.asarray(radii, dtype=np.float64).ravel()
        centers = np.asarray(centers, dtype=np.float64)
        if centers.shape != (len(radii), 2):
            raise ValueError(
                f"centers shape {centers.shape} does not match "
                f"({len(radii)}, 2) for {len(radii)} cylinders.")
        if np.any(radii <= 0):
            raise ValueError("All radii must be positive.")
        self.surface_relaxivity_t2 = _opt_float(surface_relaxivity_t2)
        self.permeability = _opt_float(permeability)
        self._L_float = float(L)
        self._radii_np = radii.copy()
        self._centers_np = centers.copy()
        o = np.asarray(orientation, dtype=np.float64)
        self.orientation = (o / np.linalg.norm(o)).astype(np.float32)
        R = _rotation_to_z(self.orientation)
        self._R = R.astype(np.float32)
        self._R_inv = R.T.astype(np.float32)
        self._is_identity_rotation = bool(np.allclose(R, np.eye(3)))
        self.min_gap = _min_gap(centers, radii, float(L))

    def init_positions(self, n_walkers, key=0):
        rng = _rng_from_key(key)
        L = self._L_float
        out, have = [], 0
        while have < n_walkers:
            xy = rng.uniform(-L / 2.0, L / 2.0, (max(n_walkers * 4, 1024), 2))
            xy = xy[_outside_all(xy, self._centers_np, self._radii_np, L)]
            out.append(xy)
            have += len(xy)
        xy = np.concatenate(out)[:n_walkers].astype(np.float32)
        r_cyl = np.concatenate([xy, np.zeros((n_walkers, 1), np.float32)], axis=1)
        return (self._R_inv.astype(np.float64) @ r_cyl.T).T.astype(np.float32)

    def volume(self, L=1.0):
        return float(np.pi * np.sum(self._radii_np ** 2) * float(L))

    def surface_area(self, L=1.0):
        return float(2.0 * np.pi * np.sum(self._radii_np) * float(L))

    def volume_fraction(self):
        return float(np.pi * np.sum(self._radii_np ** 2) / self._L_float ** 2)


class PackedSpheres(Geometry):
    """Spheres in a periodic cubic cell of side ``L`` (3-D analogue of
    ``PackedCylinders``).  Compartments: 0 outside, k inside sphere k (1-based)."""
    _walk_kind = KIND_PACKED_SPH

    def __init__(self, radii, centers, L, surface_relaxivity_t2=None, permeability=None):
        radii = np.asarray(radii, dtype=np.float64).ravel()
        centers = np.asarray(centers, dtype=np.float64)
        if centers.shape != (len(radii), 3):
            raise ValueError(
                f"centers shape {centers.shape} does not match "
                f"({len(radii)}, 3) for {len(radii)} spheres.")
        if np.any(radii <= 0):
            raise ValueError("All radii must be positive.")
        self.surface_relaxivity_t2 = _opt_float(surface_relaxivity_t2)
        self.permeability = _opt_float(permeability)
        self._L_float = float(L)
        self._radii_np = radii.copy()
        self._centers_np = centers.copy()
        self.min_gap = _min_gap(centers, radii, float(L))

    def init_positions(self, n_walkers, key=0):
        rng = _rng_from_key(key)
        L = self._L_float
        out, have = [], 0
        while have < n_walkers:
            pts = rng.uniform(-L / 2.0, L / 2.0, (max(n_walkers * 4, 1024), 3))
            pts = pts[_outside_all(pts, self._centers_np, self._radii_np, L)]
            out.append(pts)
            have += len(pts)
        return np.concatenate(out)[:n_walkers].astype(np.float32)

    def volume(self):
        return float((4.0 / 3.0) * np.pi * np.sum(self._radii_np ** 3))

    def surface_area(self):
        return float(4.0 * np.pi * np.sum(self._radii_np ** 2))

    def volume_fraction(self):
        return float((4.0 / 3.0) * np.pi * np.sum(self._radii_np ** 3) / self._L_float ** 3)


class MyelinatedCylinder(Geometry):
    """Three concentric compartments along ``orientation``: 0 intra-axonal
    (r < inner_radius), 1 myelin sheath, 2 extra-axonal (r >= outer_radius), each
    with its own isotropic diffusivity and optional T2, and a permeability per
    membrane (dmipy_sim/geometries.py:847-1136; step physics.py:420-679).

    ``water_fractions`` are per-compartment proton densities used as ensemble
    spin weights of each walker's start compartment; the default is
    (1, MYELIN_WATER_PROTON_DENSITY, 1).  ``D_myelin`` defaults to 0 (a stuck
    pool).  ``simulate`` takes no diffusivity for this geometry, ignores its T2
    argument (T2 is per compartment here) and walks one step per sample.
    """
    _walk_kind = KIND_MYELIN
    _is_myelinated = True

    def __init__(self, inner_radius, outer_radius, orientation,
                 D_intra, D_extra, D_myelin=0.0,
                 kappa_inner=None, kappa_outer=None,
                 T2_intra=None, T2_myelin=None, T2_extra=None,
                 water_fractions=None):
        if outer_radius <= inner_radius:
            raise ValueError("outer_radius must be > inner_radius")
        self.inner_radius = float(inner_radius)
        self.outer_radius = float(outer_radius)
        self.D_intra = float(D_intra)
        self.D_myelin = float(D_myelin)
        self.D_extra = float(D_extra)
        self.kappa_inner = _opt_float(kappa_inner)
        self.kappa_outer = _opt_float(kappa_outer)
        self.T2_intra = _opt_float(T2_intra)
        self.T2_myelin = _opt_float(T2_myelin)
        self.T2_extra = _opt_float(T2_extra)
        if water_fractions is None:
            water_fractions = (1.0, MYELIN_WATER_PROTON_DENSITY, 1.0)
        self.water_fractions = tuple(float(w) for w in water_fractions)
        o = np.asarray(orientation, dtype=np.float64)
        self.orientation = (o / np.linalg.norm(o)).astype(np.float32)
        R = _rotation_to_z(self.orientation)
        self._R = R.astype(np.float32)
        self._R_inv = R.T.astype(np.float32)
        self._is_identity_rotation = bool(np.allclose(R, np.eye(3)))

    def _areas(self):
        """Cross-section areas of the three compartments, the extra-axonal one
        bounded at 2 R_out (the placement convention of the reference)."""
        Ri, Ro = self.inner_radius, self.outer_radius
        return (np.pi * Ri ** 2, np.pi * (Ro ** 2 - Ri ** 2), np.pi * ((2.0 * Ro) ** 2 - Ro ** 2))

    def start_counts(self, n_walkers):
        """Walkers starting in each compartment, allotted by area in index order:
        round(N a_c / sum a) for intra and sheath, the rest extra-axonal."""
        a = self._areas()
        tot = a[0] + a[1] + a[2]
        n_intra = int(round(n_walkers * a[0] / tot))
        n_myelin = int(round(n_walkers * a[1] / tot))
        return n_intra, n_myelin, n_walkers - n_intra - n_myelin

    def init_positions(self, n_walkers, key=0):
        """Host placement (uniform in each compartment's cross-section, z = 0, lab
        frame); also sets ``_init_compartments``.  The GPU walk places walkers
        itself with the same allotment."""
        rng = _rng_from_key(key)
        counts = self.start_counts(n_walkers)
        Ri, Ro = self.inner_radius, self.outer_radius
        bands = ((0.0, Ri), (Ri, Ro), (Ro, 2.0 * Ro))
        pos = np.zeros((n_walkers, 3), np.float32)
        comp = np.zeros(n_walkers, np.int32)
        i = 0
        for c, (n_c, (lo, hi)) in enumerate(zip(counts, bands)):
            got, have = [], 0
            while have < n_c:
                xy = rng.uniform(-hi, hi, (n_c * 3, 2))
                r = np.linalg.norm(xy, axis=1)
                xy = xy[(r >= lo) & (r < hi)]
                got.append(xy)
                have += len(xy)
            if n_c:
                pos[i:i + n_c, :2] = np.concatenate(got)[:n_c]
                comp[i:i + n_c] = c
            i += n_c
        self._init_compartments = comp
        return (self._R_inv.astype(np.float64) @ pos.T.astype(np.float64)).T.astype(np.float32)

    def classify_position(self, r):
        """0 intra, 1 sheath, 2 extra by the squared axial-plane distance in float32."""
        r = np.asarray(r, dtype=np.float32)
        rc = r if self._is_identity_rotation else (r.astype(np.float64) @ self._R.T.astype(np.float64)).astype(np.float32)
        rr = rc[..., 0] * rc[..., 0] + rc[..., 1] * rc[..., 1]
        Ri, Ro = np.float32(self.inner_radius), np.float32(self.outer_radius)
        return np.where(rr < Ri * Ri, 0, np.where(rr < Ro * Ro, 1, 2)).astype(np.int32)

    def volume(self, compartment, L=1.0):
        names = ('intra', 'myelin', 'extra')
        if compartment not in names:
            raise ValueError(
                f"compartment must be 'intra', 'myelin', or 'extra'; got {compartment!r}")
        return self._areas()[names.index(compartment)] * float(L)

    def surface_area(self, compartment, L=1.0):
        Ri, Ro = self.inner_radius, self.outer_radius
        per = {'intra': Ri, 'myelin': Ri + Ro, 'extra': Ro}
        if compartment not in per:
            raise ValueError(
                f"compartment must be 'intra', 'myelin', or 'extra'; got {compartment!r}")
        return 2.0 * np.pi * per[compartment] * float(L)

    def volume_fraction(self, compartment):
        return self.volume(compartment) / (np.pi * (2.0 * self.outer_radius) ** 2)


class Ellipsoid(Geometry):
    """Axis-aligned ellipsoid x^2/a^2 + y^2/b^2 + z^2/c^2 = 1, reflecting,
    relaxing or permeable (dmipy_sim/geometries.py:1137-1425)."""
    _walk_kind = KIND_ELLIPSOID

    def __init__(self, semiaxes, surface_relaxivity_t2=None, permeability=None):
        self.semiaxes = np.asarray(semiaxes, dtype=np.float64)
        if self.semiaxes.shape != (3,) or np.any(self.semiaxes <= 0):
            raise ValueError("semiaxes must be three positive lengths")
        self.surface_relaxivity_t2 = _opt_float(surface_relaxivity_t2)
        self.permeability = _opt_float(permeability)

    def init_positions(self, n_walkers, key=0):
        """Uniform in the ellipsoid: unit-ball rejection scaled by the semi-axes."""
        rng = _rng_from_key(key)
        out, have = [], 0
        while have < n_walkers:
            pts = rng.uniform(-1.0, 1.0, (n_walkers * 4, 3))
            pts = pts[np.linalg.norm(pts, axis=1) < 1.0]
            out.append(pts)
            have += len(pts)
        return (np.concatenate(out)[:n_walkers] * self.semiaxes).astype(np.float32)

    def volume(self):
        a, b, c = self.semiaxes
        return (4.0 / 3.0) * np.pi * a * b * c

    def surface_area(self):
        """Thomsen's approximation, p = 1.6075 (relative error < 1.061 %)."""
        p = 1.6075
        ap, bp, cp = self.semiaxes ** p
        return 4.0 * np.pi * ((ap * bp + ap * cp + bp * cp) / 3.0) ** (1.0 / p)


class PermeableSlab1D(Geometry):
    """Closed 1-D two-compartment slab: permeable membrane at x = L/2, reflecting
    walls at 0 and L, y and z free (dmipy_sim/geometries.py:3221-3300).  Walkers
    start in compartment A = {x < L/2}; f_A(t) = 1/2 + 1/2 exp(-4 kappa t / L)."""
    _walk_kind = KIND_SLAB

    def __init__(self, length, permeability, surface_relaxivity_t2=None):
        self.length = float(length)
        self.permeability = float(permeability)
        self.surface_relaxivity_t2 = _opt_float(surface_relaxivity_t2)
        self.radius = float(length) / 2.0        # sub-step length scale

    def volume(self):
        return self.length / 2.0

    def surface_area(self):
        return 1.0

    def init_positions(self, n_walkers, key=0):
        rng = _rng_from_key(key)
        x = rng.uniform(0.0, self.length / 2.0, n_walkers)
        z = np.zeros(n_walkers)
        return np.stack([x, z, z], axis=1).astype(np.float32)


class PermeableShell(Geometry):
    """Closed radial two-compartment shell: permeable membrane at r = R_in inside a
    reflecting wall at r = R_out; ``kind='sphere'`` (r = |x|) or ``'cylinder'``
    (r = distance from the axis along ``orientation``)
    (dmipy_sim/geometries.py:3303-3431).  The cylindrical shell walks in its
    native frame (axis = z)."""

    def __init__(self, r_inner, r_outer, permeability, kind='sphere',
                 orientation=(0.0, 0.0, 1.0), surface_relaxivity_t2=None):
        if kind not in ('sphere', 'cylinder'):
            raise ValueError("kind must be 'sphere' or 'cylinder'")
        self.r_inner = float(r_inner)
        self.r_outer = float(r_outer)
        self.permeability = float(permeability)
        self.kind = kind
        self.surface_relaxivity_t2 = _opt_float(surface_relaxivity_t2)
        self.radius = float(r_inner)              # sub-step length scale
        o = np.asarray(orientation, dtype=np.float64)
        self._o = o / np.linalg.norm(o)
        self._walk_kind = KIND_SHELL if kind == 'sphere' else KIND_SHELL_CYL
        if kind == 'cylinder':
            R = _rotation_to_z(self._o.astype(np.float32))
            self._R = R.astype(np.float32)
            self._R_inv = R.T.astype(np.float32)
            self._is_identity_rotation = bool(np.allclose(R, np.eye(3)))

    def volume(self):
        if self.kind == 'sphere':
            return (4.0 / 3.0) * np.pi * self.r_inner ** 3
        return np.pi * self.r_inner ** 2

    def surface_area(self):
        if self.kind == 'sphere':
            return 4.0 * np.pi * self.r_inner ** 2
        return 2.0 * np.pi * self.r_inner

    def init_positions(self, n_walkers, key=0):
        """Uniform in compartment A (r < R_in); the cylindrical shell also spreads
        walkers over [0, R_in) along its axis."""
        rng = _rng_from_key(key)
        if self.kind == 'sphere':
            return Sphere(self.r_inner).init_positions(n_walkers, key)
        out, have = [], 0
        while have < n_walkers:
            xy = rng.uniform(-self.r_inner, self.r_inner, (n_walkers * 2, 2))
            out.append(xy[np.linalg.norm(xy, axis=1) < self.r_inner])
            have += len(out[-1])
        xy = np.concatenate(out)[:n_walkers]
        zc = rng.uniform(0.0, self.r_inner, n_walkers)
        native = np.column_stack([xy, zc])
        if getattr(self, '_is_identity_rotation', True):
            return native.astype(np.float32)
        return (native @ self._R.astype(np.float64)).astype(np.float32)
