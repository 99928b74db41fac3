"""Benchmark of the batched Cholesky path on one MI355X.

Workload: the 15 ranked (batch, n) shapes of the GPU MODE cholesky task
(4096x32 ... 1x32768).  Inputs are synthetic and seeded, generated on the GPU
and resident in HBM before timing:
  n = 32          dense  G G^T/n + 1.0 I
  64 <= n <= 1024 spectral Q diag(logspace(0,-3,n)) Q^T (cond 1e3)
  n >= 2048       dense  G G^T/n + 0.1 I
A step is one pass over the 15 shapes.  Per shape, each step flushes 512 MiB
of cache and times, with HIP events, one call per rotated input
(min(50, 256 MiB / input) inputs, as the reference's eval loop does); the
per-call mean over the K steps is that shape's time.  The metric is the
geometric mean of the 15 per-shape means (microseconds, lower is better).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--mode plain|mixed|both]

Both forms are timed by default; the headline value is the mixed form
(BF16x3 trailing updates), reported with its FP64 residual beside the plain
all-FP32 form's.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import torch

from gpumode_cholesky_amd import cholesky
from gpumode_cholesky_amd import harness

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3


def _parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--mode", default="both", choices=["plain", "mixed", "both"])
    ap.add_argument("--shapes", default="", help="comma list like 4096x32,1x4096 (default: all 15)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget-s", type=float, default=15.0)
    ap.add_argument("--detail", default="", help="write per-shape JSON here")
    return ap.parse_args()


def _specs(filter_s: str):
    specs = harness.BENCH_SPECS
    if filter_s:
        want = {tuple(int(x) for x in s.split("x")) for s in filter_s.split(",") if s}
        specs = [s for s in specs if (s["batch"], s["n"]) in want]
    return specs


def cpu_baseline(budget_s: float):
    """Time the FP64 C oracle (the CPU port of the path) on the host cores.

    Sample: per shape, as many matrices of that shape as fit a share of the
    budget (at least one) for n <= 2048; n >= 4096 is extrapolated from one
    n = 2048 matrix by (n/2048)^3.  Per-shape time = per-matrix time x batch;
    the value is the geometric mean over the 15 shapes, as for the GPU.
    """
    import numpy as np

    import oracle

    oracle.build()
    threads = os.cpu_count() or 1
    per_matrix = {}
    sample = []
    share = budget_s / 10.0
    for spec in harness.BENCH_SPECS:
        b, n = spec["batch"], spec["n"]
        if n > 2048 or n in per_matrix:
            continue
        a = harness.config_input(1, n, spec["seed"], device="cpu").numpy()
        t0 = time.perf_counter()
        oracle.potrf64(a)
        one = time.perf_counter() - t0
        k = int(max(1, min(b, share / max(one, 1e-6))))
        batch_a = np.repeat(a, k, axis=0)
        t0 = time.perf_counter()
        oracle.potrf64(batch_a)
        per_matrix[n] = (time.perf_counter() - t0) / k
        sample.append(f"{k}x{n}")
    shape_us = []
    for spec in harness.BENCH_SPECS:
        b, n = spec["batch"], spec["n"]
        t = per_matrix[n] if n in per_matrix else per_matrix[2048] * (n / 2048.0) ** 3
        shape_us.append(t * b * 1e6)
    geo = math.exp(sum(math.log(x) for x in shape_us) / len(shape_us))
    return {
        "value": geo,
        "unit": "us (geomean over 15 shapes)",
        "cores": threads,
        "kind": "port",
        "sample": ("FP64 C oracle (oracle/cholesky_oracle.c, OpenMP over the batch) on "
                   + ", ".join(sample) + " matrices; per-shape time = per-matrix time x batch; "
                   "n >= 4096 extrapolated from n = 2048 by n^3"),
        "per_shape_us": [round(x, 1) for x in shape_us],
    }


def _inputs(specs, dev):
    data = {}
    for spec in specs:
        b, n = spec["batch"], spec["n"]
        data[(b, n)] = [harness.config_input(b, n, spec["seed"] + 42 * i, device=dev)
                        for i in range(harness.rotation_count(b, n))]
        print(f"# inputs {b}x{n} x{len(data[(b, n)])}", file=sys.stderr, flush=True)
    return data


def _gate(specs, data, mode):
    """Correctness of one form on every shape (untimed): the reference's gate
    and the FP64 residual ||LL^T - A||_F / ||A||_F."""
    out = {}
    for spec in specs:
        key = (spec["batch"], spec["n"])
        a = data[key][0]
        L = cholesky(a, mode=mode)
        torch.cuda.synchronize()
        ok, msg = harness.check_implementation(a, L)
        if not ok:
            print(json.dumps({"error": f"{mode} {key[0]}x{key[1]}: {msg}"}))
            sys.exit(1)
        out[key] = harness.fp64_residual(a, L)
        del L
        print(f"# gate {mode} {key[0]}x{key[1]}: {msg} fp64_resid={out[key]:.2e}",
              file=sys.stderr, flush=True)
    return out


def _time(specs, data, mode, steps, warmup, dev):
    fn = (lambda d: cholesky(d, mode=mode))
    for _ in range(warmup):
        for spec in specs:
            for d in data[(spec["batch"], spec["n"])]:
                fn(d)
    torch.cuda.synchronize()
    ev = {(s["batch"], s["n"]): [] for s in specs}
    for step in range(steps):
        for spec in specs:
            key = (spec["batch"], spec["n"])
            harness.flush_caches(dev)
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            for d in data[key]:
                fn(d)
            e1.record()
            ev[key].append((e0, e1, len(data[key])))
        torch.cuda.synchronize()
        print(f"# {mode} step {step + 1}/{steps} done", file=sys.stderr, flush=True)
    us = {}
    for key, lst in ev.items():
        vals = [e0.elapsed_time(e1) * 1e3 / c for (e0, e1, c) in lst]
        us[key] = sum(vals) / len(vals)
    return us


def _geo(xs):
    xs = list(xs)
    return math.exp(sum(math.log(x) for x in xs) / len(xs))


def main():
    args = _parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; --gpus must be 1 "
                          "(a multi-GPU run is not part of this build)"}))
        sys.exit(2)
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    specs = _specs(args.shapes)
    modes = ["plain", "mixed"] if args.mode == "both" else [args.mode]
    data = _inputs(specs, dev)
    resid = {m: _gate(specs, data, m) for m in modes}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    times = {m: _time(specs, data, m, args.steps, args.warmup, dev) for m in modes}
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0

    # The headline form: mixed when it was run (BF16x3 trailing updates), else plain.
    head = "mixed" if "mixed" in modes else modes[0]
    rows = []
    for spec in specs:
        key = (spec["batch"], spec["n"])
        b, n = key
        nbytes = 8.0 * b * n * n
        flops = b * n ** 3 / 3.0
        row = {"batch": b, "n": n}
        for m in modes:
            us = times[m][key]
            row[m] = {"us": round(us, 2), "GBps": round(nbytes / us / 1e3, 1),
                      "TFLOPs": round(flops / us / 1e6, 3),
                      "hbm_frac": round(nbytes / us / 1e3 / HBM_PEAK_GBS, 4),
                      "fp32_frac": round(flops / us / 1e6 / FP32_PEAK_TFLOPS, 4),
                      "fp64_resid": resid[m][key]}
        rows.append(row)
    geo = {m: _geo(times[m][k] for k in times[m]) for m in modes}
    ms_per_step = sum(times[head].values()) / 1e3

    # Headline roofline: the single-launch wave-register kernel on 4096x32
    # (one launch per call, so the per-call event time is its duration).
    roof = None
    r32 = next((r for r in rows if (r["batch"], r["n"]) == (4096, 32)), None)
    if r32 is not None:
        g = r32[head]["GBps"]
        roof = {"kernel": "k_potrf_small<32,vec> on 4096x32", "bound": "hbm",
                "achieved": g, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(g / HBM_PEAK_GBS, 4), "traffic": None}
    out = {
        "metric": "geomean us over 15 batch x n SPD shapes",
        "value": round(geo[head], 2),
        "unit": "us",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if head == "plain" else "f32+bf16x3",
        "data": "synthetic",
        "config": {"workload": "15-shape batched Cholesky grid (4096x32 ... 1x32768), FP32 in/out",
                   "form": head, "shapes": len(rows)},
        "roofline": roof,
        "geomean_us": {m: round(geo[m], 2) for m in modes},
        "max_fp64_resid": {m: max(resid[m].values()) for m in modes},
        "wall_s_timed": round(wall, 2),
        "per_shape": rows,
    }
    if len(modes) == 2:
        out["mixed_speedup_vs_plain"] = round(geo["plain"] / geo["mixed"], 4)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.cpu_budget_s)
    if args.detail:
        with open(args.detail, "w") as f:
            json.dump(out, f, indent=1)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
