"""Offline autotuner of the launch variants (ex_tuning_*, include/exafac_hip.h).

    python -m rdist_tensor_amd.autotune [--scale S] [--out PATH] [--json PATH]

For each knob it runs STS-CP ALS sweeps on the GPU with every compiled
variant selected in turn and reads the library's own span timing (HIP events
around the launch group the knob controls): the STS-CP draw for the
sts_draw.* knobs, the sampled-MTTKRP row pass for row_scan.*.  Workloads are
the STS-CP benchmark tensors (rdist_tensor_amd/synthetic.py): 1e6 x 1e6 x 1e5
with Zipf(1.1) coordinates at R = 32 (and R = 16 for the r16 knobs), and the
four-mode 1e7 x 1e6 x 1e4 x 1e2 tensor at R = 64 for sts_draw.fp32.r64, at
`scale` of their nonzeros (the draws' cost depends on the mode sizes, which
stay full).  Every variant starts from the same factors and seed.  The
fastest mean launch wins; a variant must beat the default by 1% to replace
it.  The result is written as tuning.txt ("knob value" lines, the measured
times in comments), which the first ex_ctx_create of a process loads from
the library's directory.
"""
import argparse
import datetime
import json
import os
import sys
import time

KNOBS = {
    # knob: (workload, R, precision, timed span)
    "sts_draw.fp64.r32": ("sts_1e6x1e6x1e5", 32, "fp64", "sts_draw"),
    "sts_draw.fp32.r32": ("sts_1e6x1e6x1e5", 32, "mixed", "sts_draw"),
    "row_scan.fp64": ("sts_1e6x1e6x1e5", 32, "fp64", "row_scan"),
    "row_scan.fp32": ("sts_1e6x1e6x1e5", 32, "mixed", "row_scan"),
    "sts_draw.fp64.r16": ("sts_1e6x1e6x1e5", 16, "fp64", "sts_draw"),
    "sts_draw.fp32.r16": ("sts_1e6x1e6x1e5", 16, "mixed", "sts_draw"),
    "sts_draw.fp32.r64": ("sts_1e7x1e6x1e4x1e2", 64, "mixed", "sts_draw"),
    "sts_first.fp32": ("sts_1e6x1e6x1e5", 32, "mixed", "sts_draw"),
}
SCALE = {"sts_1e6x1e6x1e5": 1.0, "sts_1e7x1e6x1e4x1e2": 0.02}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def time_variant(abi, ctx, T, cfg, U0, R, precision, span, warm, sweeps):
    """Mean ms per launch of `span` over `sweeps` sweeps after `warm`."""
    prec = abi.EX_MIXED if precision == "mixed" else abi.EX_FP64
    F = abi.Factors(ctx, cfg["dims"], R)
    for k, u in enumerate(U0):
        F.set(k, u)
    S = abi.Sampler(ctx, F, abi.EX_SAMPLER_STS_CP, precision=prec, seed=cfg["seed"])
    A = abi.ALS(ctx, T, F, S, precision=prec)
    A.init()
    A.run(warm, cfg["J"], 10 ** 9)
    ctx.sync()
    ctx.timing(True)
    A.run(sweeps, cfg["J"], 10 ** 9)
    ctx.sync()
    t = json.loads(ctx.timing_json())
    ctx.timing(False)
    for x in (A, S, F):
        x.close()
    st = t["spans"][span]
    return st["ms"] / st["count"]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--scale", type=float, default=None,
                    help="fraction of each workload's nonzeros (default: full 1e6 tensor, 2%% of the four-mode one)")
    ap.add_argument("--knobs", default=",".join(KNOBS), help="comma-separated knobs to tune")
    ap.add_argument("--warm", type=int, default=1)
    ap.add_argument("--sweeps", type=int, default=3)
    ap.add_argument("--out", default=os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.txt"))
    ap.add_argument("--json", default=None, help="also write the measured times here")
    args = ap.parse_args()
    os.environ["EXAFAC_TUNING"] = "off"    # measure from the built-in defaults
    import torch  # noqa: F401  (the HIP runtime starts as in the library's other users)
    from rdist_tensor_amd import abi, synthetic
    knobs = [k for k in args.knobs.split(",") if k]
    info = abi.tuning()
    for k in knobs:
        if k not in KNOBS or k not in info:
            raise SystemExit(f"unknown knob {k}")
    ctx = abi.Context(0)
    device = ctx.device_name()
    results = {}
    by_workload = {}
    for k in knobs:
        by_workload.setdefault(KNOBS[k][0], []).append(k)
    for wl, ks in by_workload.items():
        scale = args.scale if args.scale is not None else SCALE[wl]
        t0 = time.time()
        cfg, idx, vals = synthetic.make(wl, scale)
        T = abi.Tensor(ctx, idx, vals, cfg["dims"])
        T.build_lookups()
        ctx.sync()
        log(f"[autotune] {wl}: {len(idx)} nonzeros ready in {time.time() - t0:.1f} s")
        del idx, vals
        for k in ks:
            _, R, precision, span = KNOBS[k]
            U0 = synthetic.factors(cfg["dims"], R, cfg["seed"])
            times = []
            for v in range(info[k][0]):
                abi.tuning_set(k, v)
                ms = time_variant(abi, ctx, T, cfg, U0, R, precision, span, args.warm, args.sweeps)
                times.append(ms)
                log(f"[autotune] {k} variant {v}: {ms:.4f} ms per launch")
            abi.tuning_set(k, 0)
            best = min(range(len(times)), key=lambda v: times[v])
            if times[best] > 0.99 * times[0]:
                best = 0
            results[k] = {"value": best, "ms": times, "workload": wl, "scale": scale, "R": R,
                          "precision": precision, "span": span}
        T.close()
    ctx.close()
    stamp = datetime.datetime.now(datetime.timezone.utc).strftime("%Y-%m-%d %H:%M UTC")
    lines = [f"# launch variants chosen by rdist_tensor_amd/autotune.py on {device}, {stamp}",
             f"# (mean ms per launch of the timed span over {args.sweeps} sweeps after {args.warm}, per variant)"]
    for k in KNOBS:
        if k in results:
            r = results[k]
            ms = " ".join(f"v{v}={t:.4f}" for v, t in enumerate(r["ms"]))
            lines.append(f"{k} {r['value']}   # {r['span']}, {r['workload']} x{r['scale']} R={r['R']} "
                         f"{r['precision']}: {ms}")
    text = "\n".join(lines) + "\n"
    with open(args.out, "w") as f:
        f.write(text)
    print(text, end="")
    if args.json:
        with open(args.json, "w") as f:
            json.dump({"device": device, "time": stamp, "knobs": results}, f, indent=1)


if __name__ == "__main__":
    main()
