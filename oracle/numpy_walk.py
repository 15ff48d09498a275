"""NumPy array-at-a-time restatement of the packed-cylinder walk (a CPU timing
baseline; test infrastructure only).

The reference runs its fused walk as jax.vmap over walkers of lax.scan over the
samples and sub-steps (dmipy_sim/core.py:700-727, physics.py:293-317): on the
CPU, XLA evaluates every operation of the step body over the whole walker array,
re-reading and re-writing the (N, n_meas) phase carry on every sub-step and
forming (N, N_obj) intermediates for the all-object wall test
(geometries.py:2067-2093).  This module does the same with NumPy arrays in
float32, which is the memory behaviour the C restatement (walk_oracle.c, one
walker at a time) does not have:

  direction   normal(3) / |normal|, times sqrt(6 D dt_sub)   physics.py:298-301
  permeate    every object tested; the first valid root      geometries.py:2018-2183
              in (EPS, |s|) from inside (exit) or outside
              (entry); Powles transmission u < min(1, 2 kappa/D d_perp); else
              mirror about the wall, nudged, wrong-side clamp
  phase       phi += gamma dt_sub (G(t) . r) per sub-step     physics.py:313
  ensemble    mean cos(phi) (rho = 0, no T2)                  core.py:426-429

Positions stay unfolded float32, as the reference keeps them
(geometries.py:1843-1849).  Random numbers come from numpy's default_rng, so only
the statistics, not the streams, agree with the GPU walk; tests/test_numpy_walk.py
holds it against the C oracle to Monte-Carlo precision.
"""
import numpy as np

f32 = np.float32


def _mi(q, L):
    return q - L * np.floor(q / L + f32(0.5))


def packed_cylinder_walk(centers, radii, L, kappa_over_D, eps, nudge, step, gamma_dt_sub, G, n_sub,
                         r0, seed=0, samples=None):
    """Walk walkers r0 (N, 3) through the first `samples` samples of G (n_meas,
    n_t, 3) with n_sub sub-steps each; returns (signal (n_meas,), final r)."""
    rng = np.random.default_rng(seed)
    C = np.asarray(centers, f32)[None, :, :]          # (1, K, 2)
    R = np.asarray(radii, f32)[None, :]               # (1, K)
    RR = R * R
    L, kD, eps, nudge, step = f32(L), f32(kappa_over_D), f32(eps), f32(nudge), f32(step)
    gdt = f32(gamma_dt_sub)
    r = np.array(r0, f32)
    N = len(r)
    n_meas, n_t, _ = G.shape
    T = n_t if samples is None else min(n_t, samples)
    phi = np.zeros((N, n_meas), f32)
    rows = np.arange(N)
    for t in range(T):
        g = np.ascontiguousarray(G[:, t, :], f32)       # (n_meas, 3)
        for _ in range(n_sub):
            nv = rng.standard_normal((N, 3)).astype(f32)
            st = nv / np.sqrt((nv * nv).sum(1, dtype=f32))[:, None] * step
            s2 = st[:, :2]
            sl = np.sqrt((s2 * s2).sum(1, dtype=f32))
            dh = s2 / np.maximum(sl, f32(1e-30))[:, None]
            q = _mi(r[:, None, :2] - C, L)                 # (N, K, 2)
            dist2 = (q * q).sum(2, dtype=f32)
            inside = dist2 < RR
            dp = (dh[:, None, :] * q).sum(2, dtype=f32)
            disc = dp * dp - (dist2 - RR)
            sq = np.sqrt(np.maximum(disc, f32(0)))
            tt = np.where(inside, -dp + sq, -dp - sq)
            valid = (disc > 0) & (tt > eps) & (tt < sl[:, None]) & (sl[:, None] > 0)
            tv = np.where(valid, tt, np.inf).astype(f32)
            k = np.argmin(tv, axis=1)
            hit = np.isfinite(tv[rows, k])
            u = rng.random(N).astype(f32)
            new = r + st
            if hit.any():
                h = np.flatnonzero(hit)
                kh = k[h]
                ts = tv[h, kh]
                qc = q[h, kh]
                Rh = R[0, kh]
                hr = qc + ts[:, None] * dh[h]
                hl = np.sqrt((hr * hr).sum(1, dtype=f32))
                n = hr / np.maximum(hl, f32(1e-30))[:, None]
                remaining = sl[h] - ts
                dperp = remaining * (np.sqrt(np.maximum(disc[h, kh], f32(0))) / Rh)
                transmit = u[h] < np.minimum(f32(1), f32(2) * kD * dperp)
                refl = h[~transmit]
                if refl.size:
                    m = ~transmit
                    nn, dd = n[m], dh[refl]
                    ins = inside[refl, kh[m]]
                    rf = dd - (f32(2) * (dd * nn).sum(1, dtype=f32))[:, None] * nn
                    rfn = np.sqrt((rf * rf).sum(1, dtype=f32))
                    rf = np.where(rfn[:, None] > 0, rf / np.maximum(rfn, f32(1e-30))[:, None], f32(0))
                    go = np.maximum(remaining[m] - nudge, f32(0))
                    sgn = np.where(ins, f32(-1), f32(1))[:, None]
                    qr = (Rh[m][:, None] * nn + nudge * sgn * nn) + rf * go[:, None]
                    qn = np.sqrt((qr * qr).sum(1, dtype=f32))
                    target = np.where(ins, Rh[m] - nudge, Rh[m] + nudge)
                    wrong = np.where(ins, qn >= Rh[m], qn <= Rh[m])
                    qr = np.where(wrong[:, None], qr * (target / np.maximum(qn, f32(1e-30)))[:, None], qr)
                    new[refl, :2] = r[refl, :2] + (qr - qc[m])
            r = new
            phi += gdt * (r @ g.T)                         # (N, n_meas) carry, every sub-step
    return np.cos(phi.astype(np.float64)).mean(0), r
