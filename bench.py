"""Benchmark: randomized CP-ALS sweeps/s on one MI355X, mixed vs FP64.

Workload (default): STS-CP on a 3-mode 1,000,000 x 1,000,000 x 100,000 tensor
with 5e7 distinct Zipf(1.1) nonzeros, lognormal values stored FP32, R = 32,
J = 65,536, seed 1.  A step is one ALS sweep (one randomized least-squares
solve per mode).  The same tensor (resident in HBM before either timed region)
is decomposed twice from the same initial factors: once on the plain FP64
path and once on the mixed-precision path (FP32 data, BF16 MFMA Grams, FP64
R x R solves), each for --steps timed sweeps after --warmup untimed sweeps
(the factors and sampler are reset to the initial state after the warm-up,
so both final fits are "after --steps sweeps").  `value` is the mixed path's
sweeps/s; the FP64 path's figures, the ratio and both final fits sit beside
it.  The timed sweeps are replayed from a HIP graph captured by one untimed
sweep (--graphs 0: enqueued step by step); the span figures (per-kernel
durations, rooflines) come from a second pass over the same sweeps, enqueued
step by step with HIP events around each span, which reaches the same fit.

Prints one JSON line (metric, value, unit, timing, config) with `roofline` (the dominant
kernel of the reported path, timed live with HIP events on the library's
stream) and `cpu_baseline` (the oracle, a bounded sample of the same
workload on the host cores).
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


def tree_nodes_per_level(n, F):
    """Node counts per level of the PartitionTree heap over ceil(n/F) leaves
    (root first; the last level may be partial)."""
    leaves = -(-n // F)
    total = 2 * leaves - 1
    out, d = [], 0
    while sum(out) < total:
        out.append(min(1 << d, total - sum(out)))
        d += 1
    return out


def sts_draw_bytes(dims, R, J, j, s=8):
    """Algorithmic (compulsory) bytes of one STS-CP draw over every non-target
    mode: each distinct tree node at most once (a level holds at most J
    distinct visited nodes; packed upper R(R+1)/2 values at s bytes, the
    eigen tree's too), each distinct leaf block of F = R factor rows at most
    once, the R x R eigenvector rows and the packed pinv (FP64), samples and
    weights written.  Re-reads of shared upper levels by many samples are
    cache traffic, not counted."""
    P = tri(R)
    total = 0
    for k, n in enumerate(dims):
        if k == j:
            continue
        for lv in tree_nodes_per_level(R, 1) + tree_nodes_per_level(n, R):
            total += min(lv, J) * P * s
        leaves = -(-n // R)
        total += min(leaves, J) * R * R * s + R * R * 8
    total += P * 8                          # packed pinv
    total += J * (len(dims) * 4 + 8)        # samples + weights written
    return total


def tree_build_bytes(dims, R, j, s=8):
    """One row-tree rebuild: the factor read once, every node written once."""
    n = dims[j]
    nodes = 2 * (-(-n // R)) - 1
    return n * R * s + nodes * tri(R) * s


def factor_update_bytes(dims, R, j, s=8):
    """U_j = Y G+ and the column norms: FP64 reads Y, writes U, reads and
    writes U again to scale (4 passes); mixed reads Y twice (its Gram, then
    the product) and writes U once (3 passes)."""
    return (4 if s == 8 else 3) * dims[j] * R * s


def mttkrp_bytes(N, R, Ju, H, T, s=8):
    """Algorithmic bytes of sampled MTTKRP + KRP gather per mode step: sample
    tuples and weights, gathered factor rows, touched nonzeros (coordinates +
    value), output rows written.  The lookup's own traffic (the row pass over
    every nonzero, fiber probes) is implementation cost and not counted.
    s = bytes per factor / value / output entry (8 FP64, 4 mixed); weights
    are FP64 on both paths."""
    return Ju * (4 * (N - 1) + 8) + Ju * (N - 1) * R * s + H * (4 * N + s) + T * R * s


def exact_mttkrp_bytes(dims, j, nnz, R, s=8):
    """Algorithmic bytes of one exact MTTKRP (mode j): every nonzero's
    coordinates and value, each non-target factor read once, the output
    written once."""
    N = len(dims)
    return nnz * (4 * N + s) + sum(dims[k] for k in range(N) if k != j) * R * s + dims[j] * R * s


def load_traffic(path):
    """Measured HBM bytes per launch from the committed rocprofv3 PMC summary
    (a JSON block in the markdown file): {"<span>/<precision>/<config>": bytes}."""
    try:
        text = open(path).read()
        block = text.split("```json", 1)[1].split("```", 1)[0]
        return json.loads(block)
    except (OSError, IndexError, ValueError):
        return {}


def run_gpu(args, cfg, ctx, T, U0, precision):
    from rdist_tensor_amd import abi
    dims = cfg["dims"]
    R, J = cfg["R"], cfg["J"]
    kind = abi.EX_SAMPLER_STS_CP if cfg["sampler"] == "sts_cp" else abi.EX_SAMPLER_CP_ARLS_LEV
    prec = abi.EX_MIXED if precision == "mixed" else abi.EX_FP64
    t0 = time.time()
    F = abi.Factors(ctx, dims, R)
    for k in range(len(dims)):
        F.set(k, U0[k])
    S = abi.Sampler(ctx, F, kind, precision=prec, seed=cfg["seed"])
    A = abi.ALS(ctx, T, F, S, precision=prec)
    A.init()
    ctx.sync()
    log(f"[bench] {precision}: setup (factors, sampler) {time.time() - t0:.1f} s")
    fit0 = A.fit()
    N = len(dims)
    for w in range(args.warmup):
        for j in range(N):
            A.step(j, J, 1000 + w)
    def reset():
        # back to the initial state: timed sweeps start where an untimed run would
        for k in range(N):
            F.set(k, U0[k])
            S.update(k)
        F.set_sigma(np.ones(R))
        ctx.sync()

    reset()
    log(f"[bench] {precision}: warmup {args.warmup} sweeps done")
    graphs = bool(args.graphs)
    if graphs:
        # the throughput pass: each sweep replayed from a captured HIP graph.
        # One untimed sweep sizes the buffers and captures the graph; the
        # factors and samplers are then reset from the host (the replays
        # read the rewritten factors), so every timed sweep is a replay.
        A.use_graph(True)
        try:
            abi.ALS.run_batch([A], 1, J)
        except abi.ExafacError as e:   # a capture the library refuses: time the step-by-step pass
            log(f"[bench] {precision}: no graph capture ({e}); timing the step-by-step pass")
            graphs = False
            A.use_graph(False)
        reset()
    if graphs:
        t0 = time.perf_counter()
        abi.ALS.run_batch([A], args.steps, J)
        ctx.sync()
        elapsed = time.perf_counter() - t0
        A.use_graph(False)
        fit_tp = A.fit()
        reset()
    # the instrumented pass: the same sweeps (sweep numbers 1..K) enqueued
    # step by step with the span events on, for the per-kernel figures; the
    # fit after every sweep and the factors at the checkpoints are taken
    # outside the spans
    ctx.timing(True)
    ctx.sync()
    fits, ckpt = [], {}
    dropped0 = S.dropped_draws()
    t0 = time.perf_counter()
    for sweep in range(1, args.steps + 1):
        for j in range(N):
            A.step(j, J, sweep)
        ctx.sync()
        if args.trajectory:
            ctx.timing(False)
            fits.append(A.fit())
            if sweep in checkpoints(args.steps):
                ckpt[sweep] = [F.get(k) * F.sigma()[None, :] for k in range(N)]
            ctx.timing(True, clear=False)
    ctx.sync()
    elapsed_inst = time.perf_counter() - t0
    timing = json.loads(ctx.timing_json())
    ctx.timing(False)
    fit1 = A.fit()
    dropped = S.dropped_draws() - dropped0
    if not graphs:
        elapsed, fit_tp = elapsed_inst, fit1
    log(f"[bench] {precision}: {args.steps} sweeps in {elapsed:.3f} s "
        f"({'graph replay' if graphs else 'eager'}; instrumented pass {elapsed_inst:.3f} s); "
        f"fit {fit0:.6f} -> {fit1:.6f}")
    del A, S, F
    return dict(elapsed=elapsed, elapsed_instrumented=elapsed_inst, timing=timing, fit0=fit0, fit1=fit1,
                passes_agree=bool(fit_tp == fit1), timed_by="graph replay" if graphs else "step by step",
                fits=fits, checkpoints=ckpt, dropped=dropped)


def checkpoints(K):
    """Sweeps at which the factors of both precisions are compared."""
    return sorted({1, 2, 5, 10, 20, K} & set(range(1, K + 1)))


def factor_errors(a, b):
    """Per checkpoint: the largest relative Frobenius difference over the modes
    between the mixed path's factors (columns scaled by sigma) and the FP64
    path's."""
    out = {}
    for sweep in sorted(set(a) & set(b)):
        out[str(sweep)] = max(float(np.linalg.norm(x - y) / max(np.linalg.norm(y), 1e-300))
                              for x, y in zip(a[sweep], b[sweep]))
    return out


def run_exact(args, cfg, ctx, T, U0, precision, nnz):
    """ExactALS (no sampling) from the same initial factors: --exact-steps
    timed sweeps; the exact MTTKRP launch group's roofline."""
    from rdist_tensor_amd import abi
    dims = cfg["dims"]
    R, N = cfg["R"], len(cfg["dims"])
    prec = abi.EX_MIXED if precision == "mixed" else abi.EX_FP64
    F = abi.Factors(ctx, dims, R)
    for k in range(N):
        F.set(k, U0[k])
    A = abi.ALS.exact(ctx, T, F, precision=prec)
    A.init()
    ctx.sync()
    fit0 = A.fit()
    ctx.timing(True)
    ctx.sync()
    t0 = time.perf_counter()
    for sweep in range(1, args.exact_steps + 1):
        for j in range(N):
            A.step(j, 0, sweep)
    ctx.sync()
    elapsed = time.perf_counter() - t0
    timing = json.loads(ctx.timing_json())
    ctx.timing(False)
    fit1 = A.fit()
    log(f"[bench] exact {precision}: {args.exact_steps} sweeps in {elapsed:.3f} s; fit {fit0:.6f} -> {fit1:.6f}")
    del A, F
    K = args.exact_steps
    vb = 4 if precision == "mixed" else 8
    st = timing["spans"]["exact_mttkrp"]
    avg = st["ms"] / st["count"] * 1e-3
    per = float(np.mean([exact_mttkrp_bytes(dims, j, nnz, R, vb) for j in range(N)]))
    achieved = per / avg / 1e9
    traffic = load_traffic(os.path.join(ROOT, "profiles", "hbm_traffic.md"))
    return dict(value=K / elapsed, ms_per_step=elapsed / K * 1e3,
                roofline={"kernel": "exact_mttkrp", "bound": "hbm", "achieved": round(achieved, 2),
                          "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                          "traffic": traffic.get(f"exact_mttkrp/{precision}/{args.config}") if args.scale == 1.0
                          else None,
                          "avg_launch_ms": round(avg * 1e3, 4),
                          "algorithmic_bytes_per_launch": int(per)},
                spans_ms_per_sweep={k: round(v["ms"] / K, 4) for k, v in timing["spans"].items()},
                fit={"initial": fit0, "final": fit1})


def summarize(args, cfg, g, vbytes, precision):
    """Throughput, dominant launch group and its roofline for one GPU run.
    Candidates: the STS-CP draw, the sampled MTTKRP + KRP gather, the row-tree
    rebuild and the factor update (HBM-bound: algorithmic bytes over the
    average launch), and computeM (one-workgroup R x R eigen / Cholesky
    solves, priced at its FP64 flops against the dense FP64 matrix peak).
    The candidate with the largest total time is the dominant kernel."""
    dims = cfg["dims"]
    N, R, J = len(dims), cfg["R"], cfg["J"]
    K = args.steps
    spans = g["timing"]["spans"]
    cnt = g["timing"]["counters"]
    steps = max(cnt.get("mode_steps", 1), 1)
    traffic = load_traffic(os.path.join(ROOT, "profiles", "hbm_traffic.md"))
    mean = lambda f: float(np.mean([f(j) for j in range(N)]))
    roofs = {}
    def per_launch(name):
        st = spans[name]
        return st["ms"], st["ms"] / st["count"] * 1e-3
    if "sts_draw" in spans:
        tot, avg = per_launch("sts_draw")
        roofs["sts_draw"] = (tot, avg, mean(lambda j: sts_draw_bytes(dims, R, J, j, s=vbytes)), "hbm")
    if "sampled_mttkrp" in spans and cnt.get("mode_steps"):
        ms = spans["sampled_mttkrp"]["ms"] + spans.get("design_gather", {"ms": 0.0})["ms"]
        per = mttkrp_bytes(N, R, cnt["J_u"] / steps, cnt["H"] / steps, cnt["T"] / steps, s=vbytes)
        roofs["sampled_mttkrp"] = (ms, ms / steps * 1e-3, per, "hbm")
    if "tree_build" in spans:
        tot, avg = per_launch("tree_build")
        roofs["tree_build"] = (tot, avg, mean(lambda j: tree_build_bytes(dims, R, j, s=vbytes)), "hbm")
    if "factor_update" in spans:
        tot, avg = per_launch("factor_update")
        roofs["factor_update"] = (tot, avg, mean(lambda j: factor_update_bytes(dims, R, j, s=vbytes)), "hbm")
    if "sts_compute_m" in spans:
        tot, avg = per_launch("sts_compute_m")
        # (N + 1) symmetric R x R factorisations (eigendecompositions or
        # Cholesky) of about R^3 flops each, plus N Hadamard products
        roofs["sts_compute_m"] = (tot, avg, float((N + 1) * R ** 3 + N * R * R), "mfma")
    out = {}
    for name, (tot, avg, per, bound) in roofs.items():
        key = f"{name}/{precision}/{args.config}"
        if bound == "hbm":
            achieved = per / avg / 1e9
            out[name] = {"kernel": name, "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                         "traffic": traffic.get(key) if args.scale == 1.0 else None,
                         "avg_launch_ms": round(avg * 1e3, 4), "algorithmic_bytes_per_launch": int(per)}
        else:
            achieved = per / avg / 1e12
            out[name] = {"kernel": name, "bound": "mfma", "achieved": round(achieved, 6),
                         "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": round(achieved / FP64_PEAK_TFLOPS, 8),
                         "traffic": None, "avg_launch_ms": round(avg * 1e3, 4),
                         "algorithmic_flops_per_launch": int(per)}
        out[name]["ms_per_sweep"] = round(tot / K, 4)
    dom = max(roofs, key=lambda n: roofs[n][0]) if roofs else None
    mt = out.get("sampled_mttkrp")
    return dict(value=K / g["elapsed"], ms_per_step=g["elapsed"] / K * 1e3,
                roofline=out.get(dom), roofline_other={k: v for k, v in out.items() if k != dom},
                sampled_mttkrp_gbs=mt["achieved"] if mt else None,
                nnz_per_s=round(cnt.get("H", 0) / g["elapsed"], 1),
                spans_ms_per_sweep={k: round(v["ms"] / K, 4) for k, v in spans.items()},
                counters_per_mode_step={k: round(v / steps, 1) for k, v in cnt.items()},
                instrumented_ms_per_step=round(g["elapsed_instrumented"] / K * 1e3, 3),
                timed_by=g["timed_by"],
                throughput_and_instrumented_fits_equal=g["passes_agree"],
                fit={"initial": g["fit0"], "final": g["fit1"]},
                fits_per_sweep=g["fits"], zero_mass_draws=g.get("dropped"))


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
    while spent < budget_s and done < 6 * N:
        t1 = time.perf_counter()
        a.step(done % N, cfg["J"], 1 + done // N)
        spent += time.perf_counter() - t1
        done += 1
        log(f"[bench] cpu baseline: mode step {done} done ({spent:.1f} s)")
    sweeps_per_s = (done / N) / spent
    return dict(value=sweeps_per_s, unit="sweeps/s", cores=threads, kind="port",
                sample=f"{done} mode steps ({done / N:.2f} sweeps) of the same workload; oracle setup "
                       f"{setup:.1f} s untimed")


def batch_sweep_bytes(batch, J, vb, counters, sweeps):
    """Algorithmic bytes of the batched sweeps, from the batch's counters:
    per mode step, the sampler state read by the draws (row, score, CDF entry
    per sample and drawn mode), the samples and weights written, the fiber
    key of each unique sample, the factor rows gathered into the design, each
    hit nonzero (row-order position, row, value), the factor rows written and
    the previous ones cleared, their scores and CDF entries, and one pass over
    the row stamps of the target mode.  Design / MTTKRP scratch round trips
    and binary-search paths are implementation traffic, not counted."""
    nz, ju, rows = counters
    total = 0.0
    for q, (dims, idx, vals, R, st) in enumerate(batch):
        N = len(dims)
        steps = sweeps * N
        total += steps * J * ((N - 1) * 20 + N * 4 + 8)
        total += ju[q] * (8 + (N - 1) * R * vb)
        total += nz[q] * (8 + vb)
        total += 2 * rows[q] * R * vb + rows[q] * 20
        total += sweeps * sum(dims) * 4
    return total


def run_batched(args):
    """Batched CP-ARLS-LEV: independent tensors (mode sizes U[500, 50000],
    nnz = 20 x the largest, R in {8, 16, 32}, J = 8192) advanced together
    by ONE kernel launch per sweep (ex_batch_*: one 1024-thread workgroup per
    tensor, every mode step inside), the launch replayed from a captured HIP
    graph.  Mixed precision is the headline; the FP64 batch and the earlier
    multi-stream runner (ex_als_run_batch, one context and stream per
    optimizer) are reported beside it; the oracle on a sample of the tensors
    is the CPU baseline.  A step is one sweep of every tensor."""
    from rdist_tensor_amd import abi, synthetic
    t0 = time.time()
    batch = synthetic.make_batch(count=args.batch_count, seed=4, scale=args.scale)
    nnz_total = sum(len(b[1]) for b in batch)
    log(f"[bench] generated {len(batch)} tensors, {nnz_total} nonzeros in {time.time() - t0:.1f} s")
    J = 8192
    ctx = abi.Context(0)
    results = {}
    for prec in ("mixed", "fp64"):
        pc = abi.EX_MIXED if prec == "mixed" else abi.EX_FP64
        t0 = time.time()
        objs = []
        for dims, idx, vals, R, st in batch:
            T = abi.Tensor(ctx, idx, vals, dims)
            F = abi.Factors(ctx, dims, R)
            for k, u in enumerate(synthetic.factors(dims, R, st)):
                F.set(k, u)
            S = abi.Sampler(ctx, F, abi.EX_SAMPLER_CP_ARLS_LEV, precision=pc, seed=st)
            A = abi.ALS(ctx, T, F, S, precision=pc)
            A.init()
            objs.append((T, F, S, A))
        B = abi.Batch([o[3] for o in objs], J)
        B.use_graph(bool(args.batch_graphs))
        log(f"[bench] {prec}: {len(objs)} optimizers and the batch ready in {time.time() - t0:.1f} s")
        B.run(args.warmup)
        c0 = B.counters()
        sec = B.run(args.steps)
        c1 = B.counters()
        d = tuple(b - a for a, b in zip(c0, c1))
        fits = np.array([o[3].fit() for o in objs])
        bad = sum(int(not np.all(np.isfinite(o[1].get(k)))) for o in objs for k in range(3))
        log(f"[bench] batched {prec}: {args.steps} sweeps of {len(objs)} tensors in {sec:.4f} s; "
            f"mean fit {np.nanmean(fits):.6f}; factors with non-finite entries {bad}")
        results[prec] = dict(sec=sec, counters=d, fits=fits, bad=bad)
        B.close()
        for o in objs:
            for x in reversed(o):
                x.close()
    # the earlier runner: every optimizer on its own context and stream
    streams_rate = None
    if args.batch_compare > 0:
        objs = []
        for dims, idx, vals, R, st in batch:
            c = abi.Context(0)
            T = abi.Tensor(c, idx, vals, dims)
            F = abi.Factors(c, dims, R)
            for k, u in enumerate(synthetic.factors(dims, R, st)):
                F.set(k, u)
            S = abi.Sampler(c, F, abi.EX_SAMPLER_CP_ARLS_LEV, precision=abi.EX_MIXED, seed=st)
            A = abi.ALS(c, T, F, S, precision=abi.EX_MIXED)
            A.init()
            objs.append((c, T, F, S, A))
        abi.ALS.run_batch([o[4] for o in objs], 1, J)
        s2 = abi.ALS.run_batch([o[4] for o in objs], args.batch_compare, J)
        streams_rate = args.batch_compare / s2
        log(f"[bench] multi-stream runner (mixed): {args.batch_compare} sweeps in {s2:.3f} s")
        for c, *_ in objs:
            c.close()
    # CPU baseline: the oracle's ALS on every 32nd tensor, one sweep each, scaled to the batch
    cpu = None
    if not args.no_cpu:
        from oracle import oracle
        oracle.build()
        sample = batch[::32]
        spent = 0.0
        for dims, idx, vals, R, st in sample:
            a = oracle.ALS(idx, vals, dims, synthetic.factors(dims, R, st), "arls", seed=st)
            t1 = time.perf_counter()
            a.run(1, J, 10 ** 9)
            spent += time.perf_counter() - t1
            log(f"[bench] cpu baseline: tensor dims {dims} R={R} done ({spent:.3f} s)")
        est = spent * len(batch) / len(sample)
        cpu = {"value": round(1.0 / est, 6), "unit": "sweeps/s", "cores": int(oracle.lib().orc_num_threads()),
               "kind": "port",
               "sample": f"oracle ALS (FP64), one sweep of {len(sample)} of the {len(batch)} tensors (every 32nd), "
                         f"{spent:.3f} s, scaled by {len(batch)}/{len(sample)}; lookups built untimed"}
    h = results["mixed"]
    rate = args.steps / h["sec"]
    per = batch_sweep_bytes(batch, J, 4, h["counters"], args.steps) / args.steps
    avg = h["sec"] / args.steps
    achieved = per / avg / 1e9
    nz_rate = float(np.sum(h["counters"][0])) / h["sec"]
    out = {
        "metric": "batched CP-ARLS-LEV sweeps/s (one sweep = every tensor of the batch)",
        "value": round(rate, 4),
        "unit": "sweeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 / rate, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (factors, values, design, MTTKRP) + f64 (Grams, solves, scores)",
        "data": "synthetic",
        "config": {"workload": f"batched_arls: {len(batch)} 3-mode tensors, mode sizes U[500, 50000] x {args.scale}, "
                               f"nnz = 20 x largest mode ({nnz_total} in all), uniform coordinates, U(0,1) values, "
                               f"R in {{8,16,32}}, J={J}, CP-ARLS-LEV, mixed precision, seed 4, one launch per sweep"
                               f"{' (graph replay)' if args.batch_graphs else ''}"},
        "roofline": {"kernel": "batch_sweep_kernel<float>", "bound": "hbm", "achieved": round(achieved, 2),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                     "traffic": (load_traffic(os.path.join(ROOT, "profiles", "hbm_traffic.md"))
                                 .get("batch_sweep_kernel/mixed/batched_arls") if args.scale == 1.0 and args.batch_count == 256 else None),
                     "avg_launch_ms": round(avg * 1e3, 4), "algorithmic_bytes_per_launch": int(per)},
        "cpu_baseline": cpu,
        "nonzeros_iterated_per_s": round(nz_rate, 1),
        "unique_samples_per_sweep": int(np.sum(h["counters"][1]) / args.steps),
        "rows_updated_per_sweep": int(np.sum(h["counters"][2]) / args.steps),
        "mean_fit": {p: float(np.nanmean(r["fits"])) for p, r in results.items()},
        "max_fit_gap_mixed_vs_fp64": float(np.nanmax(np.abs(results["mixed"]["fits"] - results["fp64"]["fits"]))),
        "factors_nonfinite": {p: r["bad"] for p, r in results.items()},
        "fp64_sweeps_per_s": round(args.steps / results["fp64"]["sec"], 4),
        "multi_stream_runner_sweeps_per_s": round(streams_rate, 4) if streams_rate else None,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="sts_1e6x1e6x1e5")
    ap.add_argument("--scale", type=float, default=1.0, help="fraction of the configured nonzeros")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--precision", choices=["both", "mixed", "fp64"], default="both")
    ap.add_argument("--mttkrp-path", choices=["auto", "row_pass", "hit_list", "buckets"], default="auto",
                    help="how the sampled MTTKRP finds the sampled fibers' nonzeros")
    ap.add_argument("--exact-steps", type=int, default=3,
                    help="ExactALS sweeps per precision reported beside the randomized paths (0: skip)")
    ap.add_argument("--graphs", type=int, default=1,
                    help="time the sweeps replayed from a captured HIP graph (1) or enqueued step by step (0); "
                         "per-kernel figures always come from a second, instrumented eager pass")
    ap.add_argument("--batch-count", type=int, default=256, help="tensors of the batched workload")
    ap.add_argument("--trajectory", type=int, default=1,
                    help="instrumented pass: fit after every sweep and factor checkpoints (mixed vs FP64)")
    ap.add_argument("--batch-graphs", type=int, default=1,
                    help="batched workload: replay the one-launch sweep from a captured HIP graph (1) or launch it (0)")
    ap.add_argument("--batch-compare", type=int, default=2,
                    help="batched workload: sweeps timed on the multi-stream runner for comparison (0: skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)
    if args.config == "batched_arls":
        run_batched(args)
        return
    from rdist_tensor_amd import synthetic
    t0 = time.time()
    cfg, idx, vals = synthetic.make(args.config, args.scale)
    U0 = synthetic.factors(cfg["dims"], cfg["R"], cfg["seed"])
    log(f"[bench] generated {len(idx)} nonzeros in {time.time() - t0:.1f} s")
    from rdist_tensor_amd import abi
    ctx = abi.Context(0)
    dev = ctx.device_name()
    t0 = time.time()
    T = abi.Tensor(ctx, idx, vals, cfg["dims"])
    T.mttkrp_path({"auto": abi.EX_MTTKRP_AUTO, "row_pass": abi.EX_MTTKRP_ROW_PASS,
                   "hit_list": abi.EX_MTTKRP_HIT_LIST, "buckets": abi.EX_MTTKRP_BUCKETS}[args.mttkrp_path])
    T.build_lookups()
    ctx.sync()
    log(f"[bench] device {dev}; tensor upload + lookups {time.time() - t0:.1f} s")
    res, ckpts = {}, {}
    for prec in (["fp64", "mixed"] if args.precision == "both" else [args.precision]):
        g = run_gpu(args, cfg, ctx, T, U0, prec)
        ckpts[prec] = g.pop("checkpoints")
        res[prec] = summarize(args, cfg, g, 4 if prec == "mixed" else 8, prec)
    exact = {}
    if args.exact_steps > 0:
        for prec in (["fp64", "mixed"] if args.precision == "both" else [args.precision]):
            exact[prec] = run_exact(args, cfg, ctx, T, U0, prec, len(idx))
    N, R, J = len(cfg["dims"]), cfg["R"], cfg["J"]
    K = args.steps
    head = "mixed" if "mixed" in res else "fp64"
    h = res[head]
    cpu = None
    if not args.no_cpu and len(idx) > 100_000_000:
        cpu = {"value": None, "unit": "sweeps/s", "cores": None, "kind": "port",
               "sample": "not run: the oracle's lookups for this many nonzeros exceed a bounded sample"}
    elif not args.no_cpu:
        cpu = run_cpu_baseline(cfg, idx, vals, U0, args.cpu_budget)
    out = {
        "metric": "ALS sweeps/s + sampled-MTTKRP HBM GB/s, mixed vs FP64 GPU at equal fit",
        "value": round(h["value"], 4),
        "unit": "sweeps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": round(h["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+bf16(mfma)/f64 solves" if head == "mixed" else "f64",
        "data": "synthetic",
        "config": {"workload": f"{args.config}: {cfg['sampler']}, dims {cfg['dims']}, nnz {len(idx)}, "
                               f"Zipf {cfg['zipf']}, {cfg['values']} values, R={R}, J={J}, seed {cfg['seed']}, "
                               f"{K} sweeps per precision, timed by {h['timed_by']}",
                   "precision": head, "sampler": cfg["sampler"]},
        "sampled_mttkrp_gbs": h["sampled_mttkrp_gbs"],
        "nnz_per_s": h["nnz_per_s"],
        "fit": h["fit"],
        "roofline": h["roofline"],
        "roofline_other": h["roofline_other"],
        "cpu_baseline": cpu,
        "device": dev,
    }
    if "mixed" in res and "fp64" in res:
        fm, fd = res["mixed"]["fits_per_sweep"], res["fp64"]["fits_per_sweep"]
        out["mixed_vs_fp64"] = {"speedup": round(res["mixed"]["value"] / res["fp64"]["value"], 3),
                                "final_fit_mixed": res["mixed"]["fit"]["final"],
                                "final_fit_fp64": res["fp64"]["fit"]["final"],
                                "final_fit_abs_diff": abs(res["mixed"]["fit"]["final"] - res["fp64"]["fit"]["final"]),
                                "fit_abs_diff_per_sweep": [round(abs(a - b), 8) for a, b in zip(fm, fd)],
                                "factor_rel_err_at_sweep": factor_errors(ckpts["mixed"], ckpts["fp64"])}
    out["paths"] = res
    if exact:
        out["exact_als"] = exact
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
