"""Synthetic substrates and protocols of the benchmark workloads.

Each builder returns the inputs of one workload as plain objects of this package
(geometry, Waveform, walker count, diffusivity).  Random choices are drawn from
fixed numpy seeds so every run builds identical inputs.
"""
import numpy as np

from . import geometries as geo
from . import waveforms as wfm

D_WATER = 2e-9


def free_pgse():
    """FreeDiffusion PGSE: 50,000 walkers, D = 2e-9 m^2/s, delta = 10 ms,
    DELTA = 40 ms, n_t = 300, bvec [1,0,0], b in {0, 1e9, 2e9} s/m^2, seed 0."""
    wf = wfm.set_b(wfm.pgse(0.01, 0.04, 0.05, [[1, 0, 0]] * 3, 300), [0.0, 1e9, 2e9])
    return dict(name="free_pgse", geometry=geo.FreeDiffusion(), waveform=wf,
                n_walkers=50_000, diffusivity=D_WATER, seed=0)


def cylinder_pgse(n_walkers=1_000_000):
    """Cylinder R = 5 um along z, surface relaxivity 1e-6 m/s, 1e6 walkers,
    D = 2e-9, square-lobe PGSE delta = 10 ms, DELTA = 40 ms, n_t = 1000,
    32 Fibonacci-sphere directions x b in {0,1,2,3}e9 (128 measurements,
    direction-major)."""
    dirs = wfm.fibonacci_directions(32)
    b = np.array([0.0, 1e9, 2e9, 3e9])
    bvecs = np.repeat(dirs, len(b), axis=0)
    bvals = np.tile(b, len(dirs))
    wf = wfm.pgse(0.01, 0.04, 0.05, bvecs, 1000, slew_rate=np.inf)
    G = np.asarray(wf.G).copy()
    nz = bvals > 0
    G[nz] = wfm.set_b(wfm.Waveform(G=G[nz], dt=wf.dt, echo_idx=wf.echo_idx), bvals[nz]).G
    G[~nz] = 0.0
    wf = wfm.Waveform(G=G, dt=wf.dt, echo_idx=wf.echo_idx, rf_events=wf.rf_events)
    geom = geo.Cylinder(radius=5e-6, orientation=(0, 0, 1), surface_relaxivity_t2=1e-6)
    return dict(name="cylinder_pgse", geometry=geom, waveform=wf, n_walkers=int(n_walkers),
                diffusivity=D_WATER, seed=0, bvals=bvals, bvecs=bvecs)


def random_sine_waveform(n_meas=90, n_t=2000, T=0.05, seed=3, gmax=(0.05, 0.3)):
    """Free waveforms: per axis a sum of 4 sines with f ~ U(10, 500) Hz,
    phase ~ U(0, 2 pi), amplitude ~ U(0, 1); time mean removed (zero net moment);
    scaled so max |G| ~ U(gmax)."""
    rng = np.random.default_rng(seed)
    dt = T / (n_t - 1)
    t = np.arange(n_t) * dt
    G = np.zeros((n_meas, n_t, 3))
    for m in range(n_meas):
        for a in range(3):
            f = rng.uniform(10, 500, 4)
            ph = rng.uniform(0, 2 * np.pi, 4)
            A = rng.uniform(0, 1, 4)
            G[m, :, a] = np.sum(A[:, None] * np.sin(2 * np.pi * f[:, None] * t[None, :] + ph[:, None]), axis=0)
        G[m] -= G[m].mean(axis=0)
        G[m] *= rng.uniform(*gmax) / np.max(np.linalg.norm(G[m], axis=1))
    return wfm.Waveform(G=G.astype(np.float32), dt=float(dt), echo_idx=n_t - 1)


def packed_cylinders_substrate(seed=0, n_cyl=64, mean_r=3e-6, shape=4.0, vf=0.6,
                               permeability=1e-5, rho=None):
    """Parallel cylinders, radii ~ Gamma(shape 4, mean 3 um), non-overlapping
    periodic packing at volume fraction 0.6, permeability 1e-5 m/s."""
    rng = np.random.default_rng(seed)
    radii = rng.gamma(shape, mean_r / shape, n_cyl)
    centers, L, _ = geo.pack_dense(radii, vf, seed=seed, dim=2, gap=0.1)
    return geo.PackedCylinders(radii, centers, L, surface_relaxivity_t2=rho,
                               permeability=permeability)


def packed_cylinders_free_waveform(n_walkers=4_000_000, seed=0):
    """PackedCylinders periodic cell (64 cylinders, VF 0.6, kappa = 1e-5 m/s,
    rho = 0, D = 2e-9), 4e6 walkers, 90 free waveforms of n_t = 2000."""
    geom = packed_cylinders_substrate(seed=seed)
    wf = random_sine_waveform()
    return dict(name="packed_cylinders_free_waveform", geometry=geom, waveform=wf,
                n_walkers=int(n_walkers), diffusivity=D_WATER, seed=0)


def batch_protocol(n_t=500, T=0.04):
    """Shared 64-measurement protocol: 32 square PGSE (delta 10 ms, DELTA 30 ms,
    8 directions x b in {0.5,1,2,3}e9) + 32 cosine OGSE (f in {50,100} Hz,
    8 directions x 2 f x b in {0.5,1}e9), n_t = 500 over 40 ms."""
    dirs = wfm.fibonacci_directions(8)
    b_p = np.array([0.5e9, 1e9, 2e9, 3e9])
    pg = wfm.pgse(0.01, 0.03, 0.05, np.repeat(dirs, len(b_p), axis=0), n_t, slew_rate=np.inf)
    pg = wfm.set_b(pg, np.tile(b_p, len(dirs)))
    Gs = [np.asarray(pg.G)]
    for f in (50.0, 100.0):
        b_o = np.array([0.5e9, 1e9])
        og = wfm.ogse(f, T, 0.05, np.repeat(dirs, len(b_o), axis=0), n_t)
        og = wfm.set_b(og, np.tile(b_o, len(dirs)))
        Gs.append(np.asarray(og.G))
    return wfm.Waveform(G=np.concatenate(Gs).astype(np.float32), dt=pg.dt, echo_idx=n_t - 1)


def batch_instances(n_inst=4096, seed=4):
    """Half Sphere (even index) half Cylinder (odd, along z): R ~ U(1,10) um,
    D ~ U(0.5,3)e-9, rho log-uniform 1e-7..1e-5 m/s, kappa log-uniform
    1e-6..1e-4 m/s, walkers log-uniform 1e4..2e5."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_inst):
        R = rng.uniform(1e-6, 10e-6)
        D = rng.uniform(0.5e-9, 3e-9)
        rho = float(np.exp(rng.uniform(np.log(1e-7), np.log(1e-5))))
        kappa = float(np.exp(rng.uniform(np.log(1e-6), np.log(1e-4))))
        nw = int(round(np.exp(rng.uniform(np.log(1e4), np.log(2e5)))))
        g = (geo.Sphere(R, surface_relaxivity_t2=rho, permeability=kappa) if i % 2 == 0
             else geo.Cylinder(R, (0, 0, 1), surface_relaxivity_t2=rho, permeability=kappa))
        out.append(dict(geometry=g, diffusivity=D, n_walkers=nw,
                        params=dict(kind=i % 2, R=R, D=D, rho=rho, kappa=kappa)))
    return out
