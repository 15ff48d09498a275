"""Throughput of the Monte-Carlo walk on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
                    [--workload cylinder|free|packed|batch|surrogate]

A step is one complete walk of the workload (all walkers, all n_t x n_sub
sub-steps, all measurements, float64 ensemble reduction) with the waveform and
geometry tables already resident in HBM.  The default workload is the
Cylinder PGSE configuration (R = 5 um, rho = 1e-6 m/s, 1e6 walkers, n_t = 1000,
128 measurements).  ``batch`` is the 4096-instance Sphere/Cylinder batch in one
persistent graph-captured launch (a step = the whole batch); ``surrogate``
trains and measures the learned surrogate against the batched walk.  Prints one
JSON line on stdout; progress goes to stderr.

Roofline: the walk keeps every walker in registers, so it is bound by float32
vector (VALU) issue, not by HBM.  ``achieved`` counts algorithmic flops per
walker-step (frozen model ``falg`` below: FMA = 2, other arithmetic = 1, RNG
integer work excluded) times walker-steps per launch, divided by the walk
kernel's average duration from HIP events on the kernel's stream.
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "dmipy-sim_amd"))
sys.path.insert(0, ROOT)

METRIC = "walker-steps/s (1 GPU); signals/s surrogate vs MC at RMSE≤5e-3 S/S0"
PEAK_FP32_TFLOPS = 157.3
# Phase product of blocks of 32 or more measurements: float16 hi + lo parts of the
# displacement sums and of G on the matrix cores, float32 accumulation.
PHASE_MFMA = "fp16 hi/lo split on matrix cores, fp32 accumulate"

# HBM bytes per walk launch from rocprofv3 --pmc FETCH_SIZE and --pmc WRITE_SIZE
# runs of this script, each scaled by the counter's reading on a 256 MiB
# device-to-device copy (FETCH x 2.000, WRITE x 1.000; profiles/walk_kernel_cylinder_pgse.md).
# None where not measured.
TRAFFIC_BYTES = {
    "cylinder_pgse": 17_374_635 + 32_172_843,
    "free_pgse": None,
    "packed_cylinders_free_waveform": None,
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def falg(cfg, plan, hbar, kbar):
    """Algorithmic float32 flops per walker-step (FMA = 2, add/mul/cmp/sqrt/div = 1):
    direction 12 + move 3 + phase (6 n_meas / n_sub + 3 [n_sub > 1]) + wall + weight."""
    from dmipy_sim_amd import geometries as geo
    n_meas = cfg["waveform"].G.shape[0]
    n_sub = plan.n_sub
    f = 12 + 3 + 6.0 * n_meas / n_sub + (3 if n_sub > 1 else 0)
    kind, wall = plan.kind, plan.wall
    if kind == geo.KIND_BOX1D and wall == 1:
        f += 12
    elif kind == geo.KIND_CYLINDER:
        f += 40 if wall == 2 else 30 + 47 * hbar
    elif kind == geo.KIND_SPHERE:
        f += 45 if wall == 2 else 40 + 60 * hbar
    elif kind == geo.KIND_PACKED_CYL:
        f += 22 * kbar + 45
    elif kind == geo.KIND_PACKED_SPH:
        f += 26 * kbar + 55
    if plan.rho_over_D != 0.0 or plan.has_t2 or plan.has_t1:
        f += 3
    return f


def batch_flops(plans, walker_steps, n_meas):
    """Sum over instances of falg x walker-steps (hits not counted: the batched
    launch has no per-instance counters, so the wall term is its hit-free part)."""
    total = 0.0
    for plan, ws in zip(plans, walker_steps):
        total += falg({"waveform": _Shape(n_meas)}, plan, 0.0, 0.0) * ws
    return total


class _Shape:
    def __init__(self, n_meas):
        import numpy as np
        self.G = np.zeros((n_meas, 1, 3), np.float32)


def run_batch(args):
    """4096 Sphere/Cylinder instances (permeable, relaxing walls; uneven walker
    counts) x a shared 64-measurement PGSE/OGSE protocol, one persistent launch."""
    import dmipy_sim_amd as ds
    from dmipy_sim_amd import configs
    wf = configs.batch_protocol()
    inst = configs.batch_instances(args.n_inst)
    if args.walker_scale != 1.0:
        for s in inst:
            s["n_walkers"] = max(64, int(s["n_walkers"] * args.walker_scale))
    t0 = time.perf_counter()
    bw = ds.BatchedWalk(inst, wf, waves_per_simd=args.waves_per_simd)
    log(f"batch: {len(inst)} instances, {int(bw.n_walkers.sum())} walkers, n_t={bw.n_t}, "
        f"n_meas={bw.n_meas}, n_sub median {int(sorted(bw.n_sub)[len(inst)//2])}, "
        f"plan {time.perf_counter() - t0:.1f} s")
    plans = [ds.make_walk_plan(s["geometry"], s["diffusivity"], wf.dt) for s in inst]
    ws = [int(s["n_walkers"]) * bw.n_t * p.n_sub for s, p in zip(inst, plans)]
    sig, st = bw.run(seed=1)
    log(f"first launch: {st['total_ms']:.1f} ms, units {st['n_units']}, grid {st['grid']}")
    for i in range(args.warmup):
        bw.run(seed=2 + i)
    kms = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        kms.append(bw.run(seed=100 + i)[1]["total_ms"])
    wall = time.perf_counter() - t0
    total_ws = st["walker_steps"]
    value = total_ws * args.steps / wall
    kmean = statistics.mean(kms)
    flops = batch_flops(plans, ws, bw.n_meas)
    achieved = flops / (kmean * 1e-3) / 1e12
    log(f"timed {args.steps}: {1e3 * wall / args.steps:.1f} ms/step, graph {kmean:.1f} ms, "
        f"{value:.4e} WS/s, {achieved:.2f} TFLOP/s")
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline_batch(inst, wf, plans, args.cpu_seconds)
        cpu["gpu_over_cpu"] = round(value / cpu["value"], 1)
    roofline = {"bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS,
                "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
                "traffic": None, "kernel": "batch_kernel", "kernel_ms": round(kmean, 3),
                "flops_per_walker_step": round(flops / total_ws, 2)}
    config = {"workload": "batch_4096_sphere_cylinder", "instances": len(inst),
              "walkers": int(bw.n_walkers.sum()), "n_t": bw.n_t, "n_meas": bw.n_meas,
              "walker_steps_per_launch": total_ws, "graph": True,
              "phase": PHASE_MFMA, "ensemble_sum": "fp64"}
    line = {"metric": METRIC, "value": value, "unit": "walker-steps/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * wall / args.steps,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic", "config": config, "roofline": roofline, "cpu_baseline": cpu}
    print(json.dumps(line), flush=True)


def cpu_baseline_batch(inst, wf, plans, target_s):
    """The oracle on a sample of the batch: every 256th instance, walkers cut so
    the sample takes about target_s seconds on the host cores."""
    from oracle import oracle as orc
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    pick = list(range(0, len(inst), 256))
    ws_total = sum(wf.G.shape[1] * plans[i].n_sub for i in pick)   # per walker, all picked
    t0 = time.perf_counter()
    for i in pick:
        s = inst[i]
        orc.run_oracle(s["geometry"], s["diffusivity"], wf, threads, seed=1, threads=threads)
    probe = time.perf_counter() - t0
    n = max(threads, int(target_s / max(probe / threads, 1e-9)) // threads * threads)
    t0 = time.perf_counter()
    for i in pick:
        s = inst[i]
        orc.run_oracle(s["geometry"], s["diffusivity"], wf, n, seed=2, threads=threads)
    el = time.perf_counter() - t0
    rate = n * ws_total / el
    log(f"cpu_baseline: {len(pick)} instances x {n} walkers in {el:.2f} s -> {rate:.3e} WS/s")
    return {"value": rate, "unit": "walker-steps/s", "cores": threads, "kind": "port",
            "sample": f"{len(pick)} of {len(inst)} instances (every 256th) x {n} walkers, "
                      "all samples and sub-steps, OpenMP over walkers"}


PEAK_BF16_TFLOPS = 2500.0
PEAK_FP64_TFLOPS = 78.6


def run_replay(args):
    """The walk-once path on the Cylinder PGSE workload: the producer walk
    (float16 positions and the rho/D = 1 boundary channel, producer sub-step
    rule) and the float64 replay contraction, trajectories resident in HBM.
    A step is one producer walk plus one replay of the 128-measurement waveform."""
    import numpy as np
    import dmipy_sim_amd as ds
    from dmipy_sim_amd import configs
    cfg = configs.cylinder_pgse()
    wf, geom, D, N = cfg["waveform"], cfg["geometry"], cfg["diffusivity"], cfg["n_walkers"]
    n_t = wf.G.shape[1]
    n_meas = wf.G.shape[0]
    rho = float(geom.surface_relaxivity_t2)
    def one(seed):
        ts = ds.TrajectorySet.simulate(N, D, geom, wf.dt * (n_t - 1), wf.dt, seed=seed,
                                       save_relaxation_data=True)
        S = ts.replay(wf.G, wf.dt, surface_relaxivity=rho, D=D, meas_tiles=args.meas_tiles)
        out = (ts.walk_ms, ts.last_replay_ms, ts.sub_steps, S)
        ts.close()
        return out
    w_ms, r_ms, sub, S = one(cfg["seed"])
    log(f"replay workload: N={N} n_t={n_t} n_meas={n_meas} sub_steps={sub}; producer {w_ms:.2f} ms, "
        f"replay {r_ms:.2f} ms; S[:4]={S[:4]}")
    for i in range(args.warmup):
        one(cfg["seed"] + 1 + i)
    walks, reps = [], []
    t0 = time.perf_counter()
    for i in range(args.steps):
        w, r, _, _ = one(cfg["seed"] + 100 + i)
        walks.append(w)
        reps.append(r)
    wall = time.perf_counter() - t0
    ws = N * n_t * sub
    value = ws * args.steps / wall
    rmean = statistics.mean(reps)
    flops = 2.0 * n_meas * 3 * n_t * N
    achieved = flops / (rmean * 1e-3) / 1e12
    log(f"timed {args.steps}: {1e3 * wall / args.steps:.1f} ms/step (producer {statistics.mean(walks):.2f} ms, "
        f"replay {rmean:.2f} ms -> {achieved:.2f} TFLOP/s fp64)")
    roofline = {"bound": "mfma", "achieved": round(achieved, 3), "peak": PEAK_FP64_TFLOPS,
                "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP64_TFLOPS, 4), "traffic": None,
                "kernel": "replay_kernel", "kernel_ms": round(rmean, 4),
                "producer_ms": round(statistics.mean(walks), 3),
                "flops_per_replay": flops}
    config = {"workload": "cylinder_pgse_replay", "walkers": N, "n_t": n_t, "n_meas": n_meas,
              "sub_steps": sub, "positions": "float16", "contraction": "fp64"}
    line = {"metric": METRIC, "value": value, "unit": "walker-steps/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * wall / args.steps,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f16/f64",
            "data": "synthetic", "config": config, "roofline": roofline, "cpu_baseline": None}
    print(json.dumps(line), flush=True)


def run_surrogate(args):
    """Surrogate signals/s (4-member 8 x 512 GELU ensemble, batch 65,536, inputs
    resident) against the batched walk at the walker count that reaches the
    same RMSE bound, both on this GPU; accuracy against held-out walks with
    2e5 walkers per substrate."""
    from dmipy_sim_amd import surrogate as sg
    t0 = time.perf_counter()
    r = sg.benchmark(n_train_inst=args.surr_instances, n_walkers=args.surr_walkers,
                     epochs=args.surr_epochs, log=log)
    wall = time.perf_counter() - t0
    sur = r["surrogate"]
    roofline = {"bound": "mfma", "achieved": round(sur["tflops"], 2), "peak": PEAK_BF16_TFLOPS,
                "unit": "TFLOP/s", "frac": round(sur["tflops"] / PEAK_BF16_TFLOPS, 4),
                "traffic": None, "kernel": "gemm_fwd_big (forward chain)",
                "kernel_ms": round(sur["ms"], 4)}
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline_surrogate(r["mc"]["walkers"], args.cpu_seconds)
        cpu["gpu_surrogate_over_cpu"] = round(sur["signals_per_s"] / cpu["value"], 1)
    config = {"workload": "surrogate_packed_cylinders", "network": "8x512 GELU, bf16 weights, fp32 accumulate",
              "members": sur["members"], "inference_batch": sur["rows"],
              "training_signals": r["train"]["signals"], "training_walkers": r["train"]["walkers"],
              "epochs": r["fit"]["epochs"], "held_out_signals": r["held_out"]["signals"],
              "held_out_walkers": r["held_out"]["walkers"], "n_t": 250, "n_meas": 64}
    line = {"metric": METRIC, "value": sur["signals_per_s"], "unit": "signals/s", "n_gpus": 1,
            "steps": 1, "warmup": 1, "ms_per_step": sur["ms"], "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "bf16", "data": "synthetic",
            "config": config, "roofline": roofline, "cpu_baseline": cpu,
            "accuracy": {"rmse_surrogate": r["rmse_surrogate"], "rmse_hybrid": r["rmse_hybrid"],
                         "fallback_fraction": r["fallback_fraction"], "bound": 5e-3},
            "mc": {"walkers": r["mc"]["walkers"], "rmse": r["mc"]["rmse"],
                   "signals_per_s": r["mc"]["signals_per_s"]},
            "gain_over_mc": r["gain"], "gain_hybrid": r["gain_hybrid"],
            "train_ms_total": r["fit"]["train_ms"], "wall_s": round(wall, 1)}
    print(json.dumps(line), flush=True)


def cpu_baseline_surrogate(n_walkers, target_s):
    """The oracle walking generator substrates on the host cores: signals/s at
    the walker count the GPU walk needs for the accuracy bound."""
    import numpy as np
    from oracle import oracle as orc
    from dmipy_sim_amd import surrogate as sg
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    rng = np.random.default_rng(11)
    p = sg.sample_params(rng)
    geom = sg.build_substrate(p, seed=5)
    wf = sg.random_waveforms(rng)
    plan = orc.oracle_plan(geom, p["D"], wf.dt)
    n_t, n_meas = wf.G.shape[1], wf.G.shape[0]
    t0 = time.perf_counter()
    orc.run_oracle(geom, p["D"], wf, threads, seed=1, threads=threads, plan=plan)
    per = (time.perf_counter() - t0) / threads
    n = max(threads, int(target_s / max(per, 1e-9)) // threads * threads)
    t0 = time.perf_counter()
    orc.run_oracle(geom, p["D"], wf, n, seed=2, threads=threads, plan=plan)
    el = time.perf_counter() - t0
    ws_rate = n * n_t * plan["n_sub"] / el
    sps = ws_rate / (n_walkers * n_t * plan["n_sub"]) * n_meas
    log(f"cpu_baseline: {n} walkers x {n_t} x n_sub {plan['n_sub']} in {el:.2f} s -> "
        f"{ws_rate:.3e} WS/s = {sps:.3g} signals/s at {n_walkers} walkers")
    return {"value": sps, "unit": "signals/s", "cores": threads, "kind": "port",
            "walker_steps_per_s": ws_rate,
            "sample": f"one generator substrate (n_sub {plan['n_sub']}), {n} walkers x "
                      f"{n_t} samples x 64 measurements; signals/s at {n_walkers} walkers"}


def build_workload(name):
    from dmipy_sim_amd import configs
    if name == "cylinder":
        return configs.cylinder_pgse()
    if name == "free":
        return configs.free_pgse()
    if name == "packed":
        return configs.packed_cylinders_free_waveform()
    raise SystemExit(f"unknown workload {name!r}")


def slice_waveform(wf, n_t):
    from dmipy_sim_amd import Waveform
    return Waveform(G=wf.G[:, :n_t].copy(), dt=wf.dt, echo_idx=n_t - 1)


def cpu_baseline(cfg, target_s, reps):
    """The oracle (reference-faithful float32 restatement: fixed bounce loops,
    all objects tested, per-sub-step phase of every measurement) on the host cores,
    over a bounded walker sample of the same workload."""
    from oracle import oracle as orc
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    geom, wf, D = cfg["geometry"], cfg["waveform"], cfg["diffusivity"]
    plan = orc.oracle_plan(geom, D, wf.dt)
    n_t = wf.G.shape[1]
    probe = 64 * threads
    t0 = time.perf_counter()
    orc.run_oracle(geom, D, wf, probe, seed=1, threads=threads, plan=plan)
    dt = time.perf_counter() - t0
    per_walker = dt / probe
    n = int(max(probe, min(cfg["n_walkers"], target_s / max(per_walker, 1e-9))))
    n = (n // threads) * threads or threads
    rates = []
    for r in range(reps):
        t0 = time.perf_counter()
        orc.run_oracle(geom, D, wf, n, seed=2 + r, threads=threads, plan=plan)
        el = time.perf_counter() - t0
        rates.append(n * n_t * plan["n_sub"] / el)
        log(f"cpu_baseline rep {r}: {n} walkers in {el:.2f} s -> {rates[-1]:.3e} WS/s")
    return {"value": statistics.median(rates), "unit": "walker-steps/s", "cores": threads,
            "kind": "port",
            "sample": (f"{n} of {cfg['n_walkers']} walkers x all {n_t} samples x n_sub "
                       f"{plan['n_sub']}, median of {reps} runs, OpenMP over walkers")}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="cylinder",
                    choices=["cylinder", "free", "packed", "batch", "surrogate", "replay"])
    ap.add_argument("--n-inst", type=int, default=4096, help="batch workload: instances")
    ap.add_argument("--walker-scale", type=float, default=1.0,
                    help="batch workload: scale every instance's walker count")
    ap.add_argument("--waves-per-simd", type=int, default=0, help="batch workload: persistent waves")
    ap.add_argument("--surr-instances", type=int, default=3125,
                    help="surrogate workload: training substrates (x 64 signals)")
    ap.add_argument("--surr-walkers", type=int, default=10_000,
                    help="surrogate workload: walkers per training substrate")
    ap.add_argument("--surr-epochs", type=int, default=60)
    ap.add_argument("--meas-tiles", type=int, default=0,
                    help="replay workload: 16-measurement tiles per pass (0 = automatic)")
    ap.add_argument("--slice-t", type=int, default=0,
                    help="packed workload: walk only the first N waveform samples")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=6.0)
    ap.add_argument("--cpu-reps", type=int, default=3)
    ap.add_argument("--meas-block", type=int, default=0)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="run only a device-to-device copy kernel of BYTES bytes (HBM "
                         "counter calibration under rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        n = args.calibrate_copy // 4
        src_t = torch.ones(n, dtype=torch.float32, device="cuda")
        dst_t = torch.empty_like(src_t)
        for _ in range(3):
            dst_t.copy_(src_t)
        torch.cuda.synchronize()
        log(f"copied {args.calibrate_copy} bytes 3 times")
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(walkers shard across GPUs only in the design).", file=sys.stderr)
        sys.exit(2)

    if args.workload == "batch":
        return run_batch(args)
    if args.workload == "surrogate":
        return run_surrogate(args)
    if args.workload == "replay":
        return run_replay(args)

    import dmipy_sim_amd as ds
    cfg = build_workload(args.workload)
    wf = cfg["waveform"]
    sliced = None
    if args.workload == "packed":
        n_slice = args.slice_t or 100
        if n_slice < wf.G.shape[1]:
            wf = slice_waveform(wf, n_slice)
            sliced = n_slice
    log(f"workload {cfg['name']}: N={cfg['n_walkers']} n_meas={wf.G.shape[0]} n_t={wf.G.shape[1]}")
    pw = ds.PlannedWalk(cfg["geometry"], cfg["diffusivity"], wf, cfg["n_walkers"],
                        meas_block=args.meas_block)
    plan = pw.plan
    log(f"plan: wall={plan.wall} n_sub={plan.n_sub} kahan={pw.kahan} "
        f"walker-steps/launch={pw.walker_steps:.3e}")
    probe = pw.run(seed=cfg["seed"], counters=True)
    hbar = probe["n_hits"] / pw.walker_steps
    kbar = probe["n_cand"] / pw.walker_steps
    log(f"counters: hits/WS={hbar:.4f} candidates/WS={kbar:.3f}; signal[:4]={probe['signal'][:4]}")
    for i in range(args.warmup):
        pw.run(seed=cfg["seed"] + 1 + i)
    kms = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        kms.append(pw.run(seed=cfg["seed"] + 100 + i)["kernel_ms"])
    wall = time.perf_counter() - t0
    ms_step = 1e3 * wall / args.steps
    value = pw.walker_steps * args.steps / wall
    kmean = statistics.mean(kms)
    f = falg(cfg, plan, hbar, kbar)
    achieved = f * pw.walker_steps / (kmean * 1e-3) / 1e12
    log(f"timed {args.steps} steps: {ms_step:.3f} ms/step, kernel {kmean:.3f} ms, "
        f"{value:.4e} WS/s, F_alg={f:.1f} -> {achieved:.2f} TFLOP/s")
    traffic = TRAFFIC_BYTES.get(cfg["name"])
    roofline = {"bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS,
                "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
                "traffic": traffic, "kernel": "walk_kernel", "kernel_ms": round(kmean, 4),
                "flops_per_walker_step": round(f, 2), "hits_per_walker_step": round(hbar, 5),
                "candidates_per_walker_step": round(kbar, 4)}
    cpu = None
    if not args.no_cpu:
        ccfg = dict(cfg)
        ccfg["waveform"] = wf
        cpu = cpu_baseline(ccfg, args.cpu_seconds, args.cpu_reps)
        cpu["gpu_over_cpu"] = round(value / cpu["value"], 1)
    config = {"workload": cfg["name"], "walkers": cfg["n_walkers"], "n_t": int(wf.G.shape[1]),
              "n_meas": int(wf.G.shape[0]), "n_sub": plan.n_sub,
              "geometry": type(cfg["geometry"]).__name__, "diffusivity": cfg["diffusivity"],
              "phase": "kahan fp32 (VALU)" if pw.kahan else PHASE_MFMA, "ensemble_sum": "fp64"}
    if sliced:
        config["time_slice"] = f"first {sliced} of 2000 samples"
    line = {"metric": METRIC, "value": value, "unit": "walker-steps/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic", "config": config, "roofline": roofline, "cpu_baseline": cpu}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
