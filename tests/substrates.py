"""Small substrates and waveforms shared by the parity tests."""
import numpy as np

import dmipy_sim_amd as ds

D = 2e-9


def small_pgse(n_t=30, n_dirs=4, T=0.012, b=(0.0, 1e9, 2e9, 3e9)):
    """Square-lobe PGSE over T seconds, n_dirs Fibonacci directions x len(b)."""
    dirs = ds.fibonacci_directions(n_dirs)
    bvecs = np.repeat(dirs, len(b), axis=0)
    bvals = np.tile(np.asarray(b, float), n_dirs)
    wf = ds.pgse(T / 4, T * 3 / 4, 0.05, bvecs, n_t, slew_rate=np.inf)
    keep = bvals > 0
    G = np.zeros_like(wf.G)
    if keep.any():
        scaled = ds.set_b(ds.Waveform(G=wf.G[keep], dt=wf.dt, echo_idx=wf.echo_idx), bvals[keep])
        G[keep] = scaled.G
    return ds.Waveform(G=G, dt=wf.dt, echo_idx=wf.echo_idx)


def packed_cyl(rho=None, kappa=None, n=8, vf=0.45, seed=3):
    rng = np.random.default_rng(seed)
    radii = rng.uniform(2e-6, 3.5e-6, n)
    c, L, _ = ds.pack_dense(radii, vf, seed=seed, dim=2)
    return ds.PackedCylinders(radii, c, L, surface_relaxivity_t2=rho, permeability=kappa)


def packed_sph(rho=None, kappa=None, n=6, vf=0.25, seed=4):
    rng = np.random.default_rng(seed)
    radii = rng.uniform(2e-6, 3.5e-6, n)
    c, L, _ = ds.pack_dense(radii, vf, seed=seed, dim=3)
    return ds.PackedSpheres(radii, c, L, surface_relaxivity_t2=rho, permeability=kappa)


def tight_cell_cyl(kappa=5e-5, rho=1e-5):
    """Two permeable cylinders (3 and 2 um) in a 9-um periodic cell: the largest radius
    plus a grid cell's expanded box reaches 0.45 of the cell side, so minimum images
    take true division and the walk's screen tests the whole cell list every sub-step
    (no cached objects)."""
    radii = np.array([3.0e-6, 2.0e-6])
    centers = np.array([[0.0, 0.0], [4.5e-6, 4.5e-6]])
    return ds.PackedCylinders(radii, centers, 9.0e-6, surface_relaxivity_t2=rho, permeability=kappa)


# (name, geometry factory, T2, T1)
CASES = [
    ("free", lambda: ds.FreeDiffusion(), None, None),
    ("free_t2", lambda: ds.FreeDiffusion(), 0.05, None),
    ("box", lambda: ds.Box1D(6e-6), None, None),
    ("box_rho", lambda: ds.Box1D(6e-6, surface_relaxivity_t2=2e-5), 0.08, None),
    ("cyl", lambda: ds.Cylinder(3e-6, (0, 0, 1)), None, None),
    ("cyl_rho", lambda: ds.Cylinder(3e-6, (0, 0, 1), surface_relaxivity_t2=2e-5), None, None),
    ("cyl_perm", lambda: ds.Cylinder(3e-6, (0, 0, 1), permeability=5e-5,
                                     surface_relaxivity_t2=1e-5), None, None),
    ("sph", lambda: ds.Sphere(3e-6), None, None),
    ("sph_rho", lambda: ds.Sphere(3e-6, surface_relaxivity_t2=2e-5), 0.06, None),
    ("sph_perm", lambda: ds.Sphere(3e-6, permeability=5e-5), None, None),
    ("pcyl", lambda: packed_cyl(), None, None),
    ("pcyl_rho", lambda: packed_cyl(rho=2e-5), None, None),
    ("pcyl_perm", lambda: packed_cyl(kappa=5e-5, rho=1e-5), None, None),
    ("psph", lambda: packed_sph(), None, None),
    ("psph_rho", lambda: packed_sph(rho=2e-5), None, None),
    ("psph_perm", lambda: packed_sph(kappa=5e-5), None, None),
    ("ell", lambda: ds.Ellipsoid((3e-6, 2e-6, 4e-6)), None, None),
    ("ell_rho", lambda: ds.Ellipsoid((3e-6, 2e-6, 4e-6), surface_relaxivity_t2=2e-5), 0.07, None),
    ("ell_perm", lambda: ds.Ellipsoid((3e-6, 2e-6, 4e-6), permeability=5e-5,
                                      surface_relaxivity_t2=1e-5), None, None),
    ("slab", lambda: ds.PermeableSlab1D(8e-6, permeability=3e-5), None, None),
    ("slab_rho", lambda: ds.PermeableSlab1D(8e-6, permeability=3e-5,
                                            surface_relaxivity_t2=2e-5), None, None),
    ("shell", lambda: ds.PermeableShell(2e-6, 4e-6, 5e-5), None, None),
    ("shell_cyl", lambda: ds.PermeableShell(2e-6, 4e-6, 5e-5, kind="cylinder",
                                            surface_relaxivity_t2=1e-5), None, None),
]


# MyelinatedCylinder runs on the single-walk path only (not the batched launch);
# axial here, the rotated frame is covered in test_gpu_myelin.py.
MYELIN_CASES = [
    ("myel", lambda: ds.MyelinatedCylinder(3e-6, 5e-6, (0, 0, 1), D_intra=2e-9, D_extra=1.5e-9,
                                           D_myelin=4e-10), None, None),
    ("myel_perm", lambda: ds.MyelinatedCylinder(
        3e-6, 5e-6, (0, 0, 1), D_intra=2e-9, D_extra=2e-9, D_myelin=5e-10, kappa_inner=5e-5,
        kappa_outer=3e-5, T2_intra=0.08, T2_myelin=0.015, T2_extra=0.1), None, None),
    ("myel_outer", lambda: ds.MyelinatedCylinder(
        2e-6, 3.5e-6, (0, 0, 1), D_intra=2e-9, D_extra=2e-9, kappa_outer=1e-4,
        T2_myelin=0.02, water_fractions=(1.0, 0.3, 0.8)), None, None),
]
WALK_CASES = CASES + MYELIN_CASES


def start_positions(geom, n, seed):
    return geom.init_positions(n, seed)


def inside_starts(geom, n, seed):
    """Half the walkers inside objects, half outside (permeable packings)."""
    r = np.array(geom.init_positions(n, seed), dtype=np.float32)
    rng = np.random.default_rng(seed + 99)
    half = n // 2
    if hasattr(geom, "_centers_np"):
        k = rng.integers(0, len(geom._radii_np), half)
        dim = geom._centers_np.shape[1]
        u = rng.standard_normal((half, dim))
        u /= np.linalg.norm(u, axis=1, keepdims=True)
        rad = geom._radii_np[k] * rng.uniform(0.0, 0.95, half) ** (1.0 / dim)
        r[:half, :dim] = geom._centers_np[k] + u * rad[:, None]
    elif isinstance(geom, (ds.Cylinder, ds.Sphere)) and geom.permeability is not None:
        R = geom.radius
        far = rng.uniform(1.05 * R, 2.0 * R, n - half)
        ang = rng.uniform(0, 2 * np.pi, n - half)
        r[half:, 0] = far * np.cos(ang)
        r[half:, 1] = far * np.sin(ang)
    return r
