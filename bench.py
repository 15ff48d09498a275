"""Throughput of the Monte-Carlo walk on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
                    [--workload packed|cylinder|free|batch|surrogate|replay]

The default workload is the packed-cylinder job the walker-steps/s target is
stated on: 64 parallel cylinders (Gamma radii, mean 3 um, VF 0.6, permeability
1e-5 m/s), 4e6 walkers, 90 free waveforms of n_t = 2000 samples, 558 sub-steps
per sample (the reference's R/25 rule).  The job is one walk of all 4e6 walkers
(the reference splits it into walker batches of 5e5 only to bound its (N, n_meas)
phase carry; here each walker's Philox stream is keyed by its global index), timed
in K sample-slices: a step is one dmw_plan_advance launch that walks every walker
through n_t / K samples (n_t / K x 558 sub-steps) from the state the previous step
left in HBM, so the K timed steps are exactly one whole job.  The W warmup steps
walk slices of another walk (seed + 1), which is then abandoned.  The geometry
tables and waveform stay resident in HBM.  ``cylinder`` is the Cylinder PGSE configuration (1e6 walkers, 128
measurements, a step = the whole walk), ``batch`` the 4096-instance
Sphere/Cylinder batch in one graph launch, ``surrogate`` the learned surrogate
against the batched walk.  Prints one JSON line on stdout; progress goes to
stderr.

Roofline: the walk keeps every walker in registers, so it is bound by float32
vector (VALU) issue, not by HBM.  ``achieved`` counts algorithmic flops per
walker-step (frozen model ``falg`` below: FMA = 2, other arithmetic = 1, RNG
integer work excluded) times walker-steps per launch, divided by the walk
kernel's average duration from HIP events on the kernel's stream.

cpu_baseline: the oracle (the reference's float32 step restated operation by
operation, with its cost structure: every object tested every sub-step, the
phase of every measurement advanced every sub-step) timed on the host cores in
three builds -- the checker build, an -O3 -march=native build, and, for the
packed workload, a NumPy array-at-a-time restatement with the reference's
(walkers x objects) intermediates -- on all cores of the affinity mask, plus a
one-core figure; ``value`` is the fastest all-core variant.
"""
import argparse
import json
import os
import statistics
import sys
import time

import numpy as np

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
    # one 100-sample launch of the 4e6-walker job (profiles/walk_kernel_packed_cylinders.md)
    "packed_cylinders_free_waveform": 312_248_042_413 + 307_345_779_053,
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
        log(f"warmup {i + 1}/{args.warmup} done")
    kms = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        kms.append(bw.run(seed=100 + i)[1]["total_ms"])
        log(f"step {i + 1}/{args.steps}: {kms[-1]:.1f} ms")   # a launch takes ~25 s
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
    threads = _affinity_cores()
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
                "traffic": None, "kernel": "mlp_infer_fused<128> (all layers of a 128-row strip in LDS)",
                "kernel_ms": round(sur["ms"], 4)}
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline_surrogate(r["mc"]["walkers"], args.cpu_seconds)
        cpu["gpu_surrogate_over_cpu"] = round(sur["signals_per_s"] / cpu["value"], 1)
    config = {"workload": "surrogate_packed_cylinders", "network": "8x512 GELU, bf16 weights, fp32 accumulate",
              "members": sur["members"], "inference_batch": sur["rows"],
              "training_signals": r["train"]["signals"], "training_walkers": r["train"]["walkers"],
              "epochs": r["fit"]["epochs"], "held_out_signals": r["held_out"]["signals"],
              "held_out_walkers": r["held_out"]["walkers"], "n_t": sg.N_T, "n_meas": sg.N_MEAS,
              "substrate": f"{sg.N_CYL} packed permeable cylinders, Gamma radii clipped at "
                           f"{sg.R_MIN * 1e6:g} um, volume fraction {sg.PRIOR['vf']}"}
    line = {"metric": METRIC, "value": sur["signals_per_s"], "unit": "signals/s", "n_gpus": 1,
            "steps": 1, "warmup": 1, "ms_per_step": sur["ms"], "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "bf16", "data": "synthetic",
            "config": config, "roofline": roofline, "cpu_baseline": cpu,
            "accuracy": {"rmse_surrogate": r["rmse_surrogate"], "rmse_hybrid": r["rmse_hybrid"],
                         "fallback_fraction": r["fallback_fraction"], "bound": 5e-3},
            "mc": {"walkers": r["mc"]["walkers"], "rmse": r["mc"]["rmse"],
                   "signals_per_s": r["mc"]["signals_per_s"]},
            "gain_over_mc": r["gain"], "gain_hybrid": r["gain_hybrid"],
            "mc_screened": r["mc_screened"], "gain_vs_screened_walk": r["gain_vs_screened"],
            "train_ms_total": r["fit"]["train_ms"], "wall_s": round(wall, 1)}
    print(json.dumps(line), flush=True)


def cpu_baseline_surrogate(n_walkers, target_s):
    """The oracle walking generator substrates on the host cores: signals/s at
    the walker count the GPU walk needs for the accuracy bound."""
    import numpy as np
    from oracle import oracle as orc
    from dmipy_sim_amd import surrogate as sg
    threads = _affinity_cores()
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
                      f"{n_t} samples x {n_meas} measurements; signals/s at {n_walkers} walkers"}


def build_workload(name):
    from dmipy_sim_amd import configs
    if name == "cylinder":
        return configs.cylinder_pgse()
    if name == "free":
        return configs.free_pgse()
    if name == "packed":
        return configs.packed_cylinders_free_waveform()
    raise SystemExit(f"unknown workload {name!r}")


def _affinity_cores():
    """Host threads for the CPU baseline: every core of the affinity mask, or the
    share the job is given when OMP_NUM_THREADS sets one."""
    avail = len(os.sched_getaffinity(0))
    try:
        share = int(os.environ.get("OMP_NUM_THREADS", avail))
    except ValueError:
        share = avail
    return max(1, min(avail, share))


def _oracle_rate(orc, cfg, wf, plan, n, threads, seed, library=None):
    t0 = time.perf_counter()
    orc.run_oracle(cfg["geometry"], cfg["diffusivity"], wf, n, seed=seed, threads=threads, plan=plan,
                   library=library)
    el = time.perf_counter() - t0
    return n * wf.G.shape[1] * plan["n_sub"] / el, el


def _sized_rate(orc, cfg, wf, plan, threads, target_s, seed, library=None, reps=1):
    """Walkers sized so one run takes about target_s (every walker walks all
    samples and sub-steps); median of reps runs."""
    probe = threads
    _, el = _oracle_rate(orc, cfg, wf, plan, probe, threads, seed, library)
    rounds = max(1, int(target_s / max(el, 1e-3)))
    n = min(cfg["n_walkers"], probe * rounds)
    rates = [_oracle_rate(orc, cfg, wf, plan, n, threads, seed + 1 + r, library)[0] for r in range(reps)]
    return statistics.median(rates), n


def cpu_baseline(cfg, target_s, reps):
    """The oracle on the host cores over a bounded walker sample of the same
    workload (every sampled walker walks all samples and sub-steps): the checker
    build and an -O3 -march=native build on all cores of the affinity mask, the
    native build on one core, and for the packed workload the NumPy
    array-at-a-time restatement.  value = the fastest all-core variant."""
    import tempfile
    from oracle import oracle as orc
    cores = _affinity_cores()
    wf = cfg["waveform"]
    plan = orc.oracle_plan(cfg["geometry"], cfg["diffusivity"], wf.dt)
    n_t = wf.G.shape[1]
    out = {"unit": "walker-steps/s", "cores": cores, "cores_available": len(os.sched_getaffinity(0)),
           "kind": "port"}
    r_chk, n_chk = _sized_rate(orc, cfg, wf, plan, cores, target_s, 10, reps=reps)
    log(f"cpu_baseline checker build, {cores} threads: {n_chk} walkers -> {r_chk:.3e} WS/s")
    variants = {"checker_build_all_cores": r_chk}
    try:
        with tempfile.TemporaryDirectory() as d:
            nat = orc.native_lib(d)
            r_nat, n_nat = _sized_rate(orc, cfg, wf, plan, cores, target_s, 20, nat, reps=reps)
            r_one, n_one = _sized_rate(orc, cfg, wf, plan, 1, target_s / 2, 30, nat)
        log(f"cpu_baseline -O3 -march=native, {cores} threads: {n_nat} walkers -> {r_nat:.3e} WS/s; "
            f"1 thread: {n_one} walkers -> {r_one:.3e} WS/s")
        variants["native_O3_all_cores"] = r_nat
        out["one_core"] = r_one
        out["one_core_build"] = "gcc -O3 -march=native"
    except Exception as exc:   # no compiler on the host: the checker build stands alone
        log(f"cpu_baseline: native build unavailable ({exc})")
    if cfg["name"] == "packed_cylinders_free_waveform":
        from oracle import numpy_walk
        geom = cfg["geometry"]
        n_np, samples = 1024, 2
        r0 = np.asarray(geom.init_positions(n_np, 3), np.float32)
        t0 = time.perf_counter()
        numpy_walk.packed_cylinder_walk(geom._centers_np, geom._radii_np, plan["cell"], plan["kappa_over_D"],
                                        plan["eps"], plan["nudge"], plan["step_len"], plan["gamma_dt_sub"],
                                        np.asarray(wf.G, np.float32), plan["n_sub"], r0, seed=1,
                                        samples=samples)
        el = time.perf_counter() - t0
        r_np = n_np * samples * plan["n_sub"] / el
        log(f"cpu_baseline NumPy array-at-a-time: {n_np} walkers x {samples} samples in {el:.1f} s -> "
            f"{r_np:.3e} WS/s")
        variants["numpy_array_at_a_time"] = r_np
    best = max(variants, key=variants.get)
    out["value"] = variants[best]
    out["fastest"] = best
    out["variants"] = {k: float(f"{v:.4e}") for k, v in variants.items()}
    out["sample"] = (f"{n_chk} of {cfg['n_walkers']} walkers x all {n_t} samples x n_sub {plan['n_sub']} "
                     f"(checker build; OpenMP over walkers, {cores} threads), median of {reps}")
    return out


def run_walk(args):
    """One walk workload with its inputs resident in HBM.  The packed job is timed
    in sample-slices of one whole walk (a step = one resumed launch); the other
    walk workloads time a whole planned walk per step."""
    import dmipy_sim_amd as ds
    cfg = build_workload(args.workload)
    wf = cfg["waveform"]
    n_job = cfg["n_walkers"]
    n_meas, n_t = int(wf.G.shape[0]), int(wf.G.shape[1])
    sliced = args.workload == "packed"
    t0 = time.perf_counter()
    pw = ds.PlannedWalk(cfg["geometry"], cfg["diffusivity"], wf, n_job, meas_block=args.meas_block)
    plan = pw.plan
    ws_job = pw.walker_steps
    if sliced:
        if not 1 <= args.steps <= n_t:
            raise SystemExit(f"bench.py: --steps must be in [1, {n_t}] (a step walks n_t / steps samples)")
        bounds = [round(k * n_t / args.steps) for k in range(args.steps + 1)]
        slices = [b - a for a, b in zip(bounds, bounds[1:])]
        log(f"workload {cfg['name']}: {n_job} walkers, n_meas={n_meas} n_t={n_t} n_sub={plan.n_sub}; "
            f"one walk in {args.steps} steps of {slices[0]} samples (plan {time.perf_counter() - t0:.1f} s)")
        wseed = cfg["seed"] + 1
        for i in range(args.warmup):
            r = pw.advance(slices[0], seed=wseed)
            log(f"  warmup {i + 1}/{args.warmup}: {r['t_done']} samples, kernel {r['kernel_ms']:.1f} ms")
        pw.rewind()
        kms, ws = [], []
        t0 = time.perf_counter()
        for i, k in enumerate(slices):
            r = pw.advance(k, seed=cfg["seed"], counters=True)
            kms.append(r["kernel_ms"])
            ws.append(n_job * k * plan.n_sub)
            if (i + 1) % max(1, args.steps // 5) == 0 or i + 1 == args.steps:
                log(f"  step {i + 1}/{args.steps}: samples [{bounds[i]}, {bounds[i + 1]}), "
                    f"kernel {kms[-1]:.1f} ms")
        wall = time.perf_counter() - t0
        assert r["t_done"] == n_t, r["t_done"]
        res = r
        value = ws_job / wall
    else:
        log(f"workload {cfg['name']}: {n_job} walkers, n_meas={n_meas} n_t={n_t} n_sub={plan.n_sub}; "
            f"a step is the whole walk (plan {time.perf_counter() - t0:.1f} s)")
        res = pw.run(seed=cfg["seed"], counters=True)
        for i in range(args.warmup):
            pw.run(seed=cfg["seed"] + 1 + i)
        kms, ws = [], []
        t0 = time.perf_counter()
        for i in range(args.steps):
            kms.append(pw.run(seed=cfg["seed"] + 1 + i)["kernel_ms"])
            ws.append(ws_job)
        wall = time.perf_counter() - t0
        value = ws_job * args.steps / wall
    ms_step = 1e3 * wall / args.steps
    hbar = res["n_hits"] / ws_job
    kbar = res["n_cand"] / ws_job
    log(f"phase: {res['phase_form']}; wall events/WS={hbar:.5f} exact-test candidates/WS={kbar:.4f}; "
        f"signal[:4]={res['signal'][:4]}")
    kmean = statistics.mean(kms)
    f = falg(cfg, plan, hbar, kbar)
    # algorithmic flops of every timed launch over the launches' summed duration
    achieved = f * sum(ws) / (sum(kms) * 1e-3) / 1e12
    log(f"timed {args.steps} steps: {ms_step:.3f} ms/step, kernel {kmean:.3f} ms (min {min(kms):.3f}), "
        f"{value:.4e} WS/s, F_alg={f:.1f} -> {achieved:.2f} TFLOP/s")
    traffic = TRAFFIC_BYTES.get(cfg["name"])
    roofline = {"bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS,
                "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
                "traffic": traffic, "kernel": "walk_kernel", "kernel_ms": round(kmean, 4),
                "kernel_ms_min": round(min(kms), 4), "walker_steps_per_launch": int(statistics.mean(ws)),
                "flops_per_walker_step": round(f, 2), "wall_events_per_walker_step": round(hbar, 5),
                "exact_test_candidates_per_walker_step": round(kbar, 5)}
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(cfg, args.cpu_seconds, args.cpu_reps)
        cpu["gpu_over_cpu"] = round(value / cpu["value"], 1)
        if "one_core" in cpu:
            cpu["gpu_over_one_core"] = round(value / cpu["one_core"], 1)
    config = {"workload": cfg["name"], "walkers": n_job, "n_t": n_t, "n_meas": n_meas,
              "n_sub": plan.n_sub, "geometry": type(cfg["geometry"]).__name__,
              "diffusivity": cfg["diffusivity"], "phase": res["phase_form"], "ensemble_sum": "fp64"}
    if sliced:
        config.update({"samples_per_step": slices[0], "steps_per_job": args.steps})
    line = {"metric": METRIC, "value": value, "unit": "walker-steps/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic", "config": config, "roofline": roofline, "cpu_baseline": cpu}
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="packed",
                    choices=["cylinder", "free", "packed", "batch", "surrogate", "replay"])
    ap.add_argument("--n-inst", type=int, default=4096, help="batch workload: instances")
    ap.add_argument("--walker-scale", type=float, default=1.0,
                    help="batch workload: scale every instance's walker count")
    ap.add_argument("--waves-per-simd", type=int, default=0, help="batch workload: persistent waves")
    ap.add_argument("--surr-instances", type=int, default=2223,
                    help="surrogate workload: training substrates (x 90 signals)")
    ap.add_argument("--surr-walkers", type=int, default=10_000,
                    help="surrogate workload: walkers per training substrate")
    ap.add_argument("--surr-epochs", type=int, default=60)
    ap.add_argument("--meas-tiles", type=int, default=0,
                    help="replay workload: 16-measurement tiles per pass (0 = automatic)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=8.0)
    ap.add_argument("--cpu-reps", type=int, default=1)
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

    return run_walk(args)


if __name__ == "__main__":
    main()
