"""The learned surrogate against the method without it (the batched walk).

The full pipeline at its specified size: 2e5 training signals (2223 substrates
of 64 packed permeable cylinders x 90 free waveforms of n_t = 2000 over 25 ms,
5000 walkers each) walked by the batched launch, a 4-member ensemble of 8 x 512
GELU networks trained on the GPU by the library's own kernels, judged against
held-out walks of 2e5 walkers.  The test asserts what the surrogate is for: at
least 100x the walk's signals/s at the walker count the walk needs to reach the
same RMSE bound (5e-3 in S/S0), with the hybrid (walk wherever the ensemble
spread exceeds the bound) inside that bound; and at least 100x a conservative
reference, the screened single walk's walker-steps/s on held-out substrates
turned into signals/s at that walker count.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.timeout(1000)
def test_surrogate_gain_over_walk(gpu):
    from dmipy_sim_amd import surrogate as sg
    print("surrogate: 2223 substrates x 90 signals x 5000 walkers, 4 members 8x512, 60 epochs",
          flush=True)
    from dmipy_sim_amd import configs
    r = sg.benchmark(n_train_inst=2223, n_walkers=5000, n_held=32, held_walkers=200_000,
                     epochs=60, members=4, infer_rows=65536, seed=3,
                     log=lambda s: print(s, flush=True), keep_ensemble=True)
    ens = r.pop("ensemble")
    print({k: r[k] for k in ("rmse_surrogate", "rmse_hybrid", "fallback_fraction",
                             "fallback_instance_fraction", "gain", "gain_hybrid",
                             "gain_vs_screened")}, flush=True)
    # the packed-cylinder workload's own substrate (64 cylinders, VF 0.6, kappa 1e-5,
    # no relaxivity) under 90 free waveforms of the generator's family
    geom = configs.packed_cylinders_substrate()
    wf = configs.random_sine_waveform(T=sg.T_WAVE, seed=21)
    S_mc, _ = sg.walk_signals([dict(geometry=geom, diffusivity=2e-9)], [wf], 200_000, seed=11)
    mean, _, _ = ens.predict(sg.features(sg.instance_params(geom, 2e-9), wf))
    rmse_ws = float(np.sqrt(np.mean((mean - S_mc[0]) ** 2)))
    print(f"packed workload substrate: rmse {rmse_ws:.4f}", flush=True)
    ens.close()
    assert r["train"]["signals"] >= 200_000
    assert r["rmse_surrogate"] <= 5e-3 and r["rmse_hybrid"] <= 5e-3
    assert r["mc"]["rmse"] <= 5e-3 or r["mc"]["walkers"] == 16384
    assert r["gain"] >= 100.0
    assert r["gain_hybrid"] >= 100.0
    # also against the screened single walk's walker-step rate, a conservative reference
    assert r["gain_vs_screened"] >= 100.0
    # the ensemble spread is informative: flagged rows are a small minority
    assert r["fallback_fraction"] <= 0.05
    assert rmse_ws <= 5e-3


def test_fallback_replaces_flagged_measurements(gpu):
    """With a zero threshold every measurement is recomputed by the walk; with
    an infinite one none is."""
    from dmipy_sim_amd import surrogate as sg
    X, S, info = sg.generate(4, 4000, seed=9, chunk=4, n_meas=90, n_t=250, T=0.025)
    ens = sg.SurrogateEnsemble(2, 128, 2, max_batch=4096, seed=1)
    ens.fit(X, S, epochs=2, batch=128)
    insts = [a for a, _ in info["instances"]]
    waves = [w for _, w in info["instances"]]
    out0, flag0, t0 = sg.predict_with_fallback(ens, insts, waves, threshold=0.0, n_walkers=4000,
                                               seed=5)
    assert flag0.all() and t0["n_fallback_instances"] == 4
    # the fallback is the walk itself: same substrates, same walker count, another seed
    assert np.sqrt(np.mean((out0.reshape(-1) - S) ** 2)) < 0.02
    outi, flagi, ti = sg.predict_with_fallback(ens, insts, waves, threshold=np.inf)
    assert not flagi.any() and ti["mc_ms"] == 0.0
    mean, _, _ = ens.predict(X)
    np.testing.assert_allclose(outi.reshape(-1), mean, rtol=0, atol=0)
    ens.close()
