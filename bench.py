"""Benchmark: randomized CP-ALS sweeps/s on one MI355X.

Workload (default): STS-CP on a 3-mode 1,000,000 x 1,000,000 x 100,000 tensor
with 5e7 distinct Zipf(1.1) nonzeros, lognormal values stored FP32 (computed
in FP64 on the plain path), R = 32, J = 65,536, seed 1.  A step is one ALS
sweep (one randomized least-squares solve per mode).  Inputs are resident in
HBM before the timed region.

Prints one JSON line (the driver's contract) with `roofline` (dominant
kernel, timed live with HIP events on the library's stream) and
`cpu_baseline` (the oracle, a bounded sample of the same workload on the host
cores).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def tri(R):
    return R * (R + 1) // 2


def tree_levels(n, F):
    leaves = -(-n // F)
    lvl, cnt = 0, 1
    while cnt * 2 <= leaves:
        lvl += 1
        cnt *= 2
    return lvl


def sts_draw_bytes(dims, R, J, j, s=8):
    """Algorithmic bytes of one STS-CP draw (all non-target modes): node reads
    per sample per level (packed upper R(R+1)/2 values, before cache reuse),
    the leaf's F = R rows, the chosen rows, samples and weights."""
    P = tri(R)
    total = 0
    for k, n in enumerate(dims):
        if k == j:
            continue
        le = tree_levels(R, 1) + 2          # root + levels + tail (eigen tree)
        lr = tree_levels(n, R) + 2          # root + levels + tail (row tree)
        total += J * ((le + lr) * P * s + R * R * s + 2 * R * s)
    total += J * P * s                      # weight: h^T pinv h
    total += J * (len(dims) * 4 + 8)        # samples + weights written
    return total


def mttkrp_bytes(N, R, Ju, H, T, s=8):
    """Algorithmic bytes of sampled MTTKRP + KRP gather per mode step:
    sample tuples and weights, gathered factor rows, touched nonzeros
    (coordinates + value), output rows."""
    return Ju * (4 * (N - 1) + s) + Ju * (N - 1) * R * s + H * (4 * N + s) + T * R * s


def seg_reduce_bytes(R, Ju, H, T, s=8):
    """Algorithmic bytes of the row-owned reduction kernel: per hit its row
    key (4), packed (sample, nonzero) index (8) and value (s); each design
    row once; each touched output row once; segment offsets."""
    return H * (4 + 8 + s) + Ju * R * s + T * R * s + (T + 1) * 4


def run_gpu(args, cfg, idx, vals, U0):
    from rdist_tensor_amd import abi
    ctx = abi.Context(0)
    dev = ctx.device_name()
    dims = cfg["dims"]
    R, J = cfg["R"], cfg["J"]
    kind = abi.EX_SAMPLER_STS_CP if cfg["sampler"] == "sts_cp" else abi.EX_SAMPLER_CP_ARLS_LEV
    t0 = time.time()
    T = abi.Tensor(ctx, idx, vals, dims)
    F = abi.Factors(ctx, dims, R)
    for k in range(len(dims)):
        F.set(k, U0[k])
    S = abi.Sampler(ctx, F, kind, seed=cfg["seed"])
    A = abi.ALS(ctx, T, F, S)
    A.init()
    ctx.sync()
    log(f"[bench] device {dev}; setup (upload, lookups, sampler) {time.time() - t0:.1f} s")
    fit0 = A.fit()
    N = len(dims)
    for w in range(args.warmup):
        for j in range(N):
            A.step(j, J, 1000 + w)
    ctx.sync()
    log(f"[bench] warmup {args.warmup} sweeps done")
    ctx.timing(True)
    ctx.sync()
    t0 = time.perf_counter()
    for sweep in range(1, args.steps + 1):
        for j in range(N):
            A.step(j, J, sweep)
    ctx.sync()
    elapsed = time.perf_counter() - t0
    timing = json.loads(ctx.timing_json())
    ctx.timing(False)
    fit1 = A.fit()
    log(f"[bench] {args.steps} sweeps in {elapsed:.3f} s; fit {fit0:.6f} -> {fit1:.6f}")
    return dict(device=dev, elapsed=elapsed, timing=timing, fit0=fit0, fit1=fit1)


def run_cpu_baseline(cfg, idx, vals, U0, budget_s=30.0):
    """The oracle (C restatement, OpenMP) on the same tensor: one mode step at
    a time until about budget_s of ALS time is spent; sweeps/s = steps/N per s."""
    from oracle import oracle
    oracle.build()
    threads = oracle.lib().orc_num_threads()
    t0 = time.time()
    a = oracle.ALS(idx, vals, cfg["dims"], U0, "sts" if cfg["sampler"] == "sts_cp" else "arls", seed=cfg["seed"])
    setup = time.time() - t0
    N = len(cfg["dims"])
    done, spent = 0, 0.0
    while spent < budget_s and done < 3 * N:
        t1 = time.perf_counter()
        a.step(done % N, cfg["J"], 1 + done // N)
        spent += time.perf_counter() - t1
        done += 1
        log(f"[bench] cpu baseline: mode step {done} done ({spent:.1f} s)")
    sweeps_per_s = (done / N) / spent
    return dict(value=sweeps_per_s, unit="sweeps/s", cores=threads, kind="port",
                sample=f"{done} mode steps ({done / N:.2f} sweeps) of the same workload; oracle setup "
                       f"{setup:.1f} s untimed")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="sts_1e6x1e6x1e5")
    ap.add_argument("--scale", type=float, default=1.0, help="fraction of the configured nonzeros")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)
    from rdist_tensor_amd import synthetic
    t0 = time.time()
    cfg, idx, vals = synthetic.make(args.config, args.scale)
    U0 = synthetic.factors(cfg["dims"], cfg["R"], cfg["seed"])
    log(f"[bench] generated {len(idx)} nonzeros in {time.time() - t0:.1f} s")
    g = run_gpu(args, cfg, idx, vals, U0)
    N, R, J = len(cfg["dims"]), cfg["R"], cfg["J"]
    K = args.steps
    ms_per_step = g["elapsed"] / K * 1e3
    spans = g["timing"]["spans"]
    cnt = g["timing"]["counters"]
    # dominant kernel (launch group) by total time
    dom = max(spans.items(), key=lambda kv: kv[1]["ms"])
    name, st = dom[0], dom[1]
    avg_s = st["ms"] / st["count"] * 1e-3
    if name == "sts_draw":
        per = np.mean([sts_draw_bytes(cfg["dims"], R, J, j) for j in range(N)])
    elif name == "seg_reduce":
        per = seg_reduce_bytes(R, cnt["J_u"] / cnt["mode_steps"], cnt["H"] / cnt["mode_steps"],
                               cnt["T"] / cnt["mode_steps"])
    else:
        per = float("nan")
    achieved = per / avg_s / 1e9
    roof = {"kernel": name, "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "avg_launch_ms": round(avg_s * 1e3, 4)}
    # sampled-MTTKRP GB/s (the metric's second half)
    mt = spans.get("sampled_mttkrp")
    mt_gbs = None
    if mt and cnt.get("mode_steps"):
        per_mt = mttkrp_bytes(N, R, cnt["J_u"] / cnt["mode_steps"], cnt["H"] / cnt["mode_steps"],
                              cnt["T"] / cnt["mode_steps"])
        mt_gbs = per_mt / (mt["ms"] / mt["count"] * 1e-3) / 1e9
    cpu = None
    if not args.no_cpu:
        cpu = run_cpu_baseline(cfg, idx, vals, U0, args.cpu_budget)
    out = {
        "metric": "ALS sweeps/s + sampled-MTTKRP HBM GB/s, mixed vs FP64 GPU at equal fit",
        "value": round(K / g["elapsed"], 4),
        "unit": "sweeps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{args.config}: STS-CP plain FP64, dims {cfg['dims']}, nnz {len(idx)}, "
                               f"Zipf {cfg['zipf']}, R={R}, J={J}, seed {cfg['seed']}",
                   "precision": "fp64", "sampler": cfg["sampler"]},
        "sampled_mttkrp_gbs": None if mt_gbs is None else round(mt_gbs, 2),
        "nnz_per_s": round(cnt.get("H", 0) / g["elapsed"], 1),
        "fit": {"initial": g["fit0"], "final": g["fit1"]},
        "spans_ms_per_sweep": {k: round(v["ms"] / K, 4) for k, v in spans.items()},
        "counters_per_mode_step": {k: round(v / max(cnt.get("mode_steps", 1), 1), 1) for k, v in cnt.items()},
        "roofline": roof,
        "cpu_baseline": cpu,
        "device": g["device"],
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
