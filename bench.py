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

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--mode plain|mixed|fp8|both]

Both forms are timed by default.  The headline value is the mixed form
(``cholesky(A, mode="mixed")``: BF16x3 panel and trailing products), the
form the metric's mixed-vs-FP32 speedup is about; the drop-in
``custom_kernel`` runs the plain all-FP32 form, whose geomean is reported
beside it (``geomean_us.plain``), with each form's FP64 residual.
``--mode fp8`` times the fp8 form alone (E4M3 bulk updates from n = 16384
on, residual ~1e-2; the extras compare it with the mixed form on the two
largest shapes in every default run).

    python bench.py --paired     # A-B-B-A plain vs mixed per shape (one JSON line)
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np
import torch

from gpumode_cholesky_amd import cholesky
from gpumode_cholesky_amd import harness

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
PMC_TRAFFIC_4096x32 = (8260.1 * 2.0 + 16384.0) * 1024  # bytes per launch


def _parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--mode", default="both", choices=["plain", "mixed", "fp8", "both"])
    ap.add_argument("--shapes", default="", help="comma list like 4096x32,1x4096 (default: all 15)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--paired", action="store_true",
                    help="same-process A-B-B-A timing of plain (A) vs mixed (B) per shape instead")
    ap.add_argument("--detail", default="", help="write per-shape JSON here")
    ap.add_argument("--no-extras", action="store_true", help="skip ragged / graph / backward measurements")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="MiB: run one device copy of this size first (scales the PMC byte counters)")
    return ap.parse_args()


def _specs(filter_s: str):
    specs = harness.BENCH_SPECS
    if filter_s:
        want = {tuple(int(x) for x in s.split("x")) for s in filter_s.split(",") if s}
        specs = [s for s in specs if (s["batch"], s["n"]) in want]
    return specs


def cpu_baseline(specs, data, host_L):
    """torch.linalg.cholesky in FP32 on the host cores: the reference's own
    path on a machine without a GPU (custom_kernel falls through to
    torch.linalg.cholesky_ex there, submission.py:4109, :4268).

    Same inputs as the GPU run (input 0 of each shape, copied to the host),
    every shape measured: median of 5 runs after a warmup for n <= 2048, of 3
    for n <= 8192, one run above.  Also gives, per shape and form, the
    max-norm forward error of the GPU factor against this one.
    """
    threads = os.cpu_count() or 1
    torch.set_num_threads(threads)
    shape_us, fwd, sample = [], {}, []
    for spec in specs:
        b, n = spec["batch"], spec["n"]
        a = data[(b, n)][0].cpu()
        reps = 5 if n <= 2048 else (3 if n <= 8192 else 1)
        if n <= 2048:
            torch.linalg.cholesky(a)
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            Lr = torch.linalg.cholesky(a)
            ts.append(time.perf_counter() - t0)
        ts.sort()
        us = ts[len(ts) // 2] * 1e6
        shape_us.append(us)
        sample.append(f"{b}x{n}:{reps}")
        scale = Lr.abs().max().item()
        fwd[(b, n)] = {m: (L - Lr).abs().max().item() / scale for m, L in host_L.get((b, n), {}).items()}
        print(f"# cpu {b}x{n}: {us:.0f} us (median of {reps}) fwd_vs_torch_cpu "
              + " ".join(f"{m}={e:.2e}" for m, e in fwd[(b, n)].items()), file=sys.stderr, flush=True)
        del a, Lr
    geo = math.exp(sum(math.log(x) for x in shape_us) / len(shape_us))
    out = {
        "value": geo,
        "unit": "us (geomean over the shapes)",
        "cores": threads,
        "kind": "reference",
        "sample": ("torch.linalg.cholesky FP32 (LAPACK spotrf via torch, the reference's custom_kernel on a "
                   "machine without a GPU) on input 0 of every shape, all measured: " + ", ".join(sample)
                   + " runs (median)"),
        "per_shape_us": [round(x, 1) for x in shape_us],
    }
    return out, fwd


def _inputs(specs, dev):
    data = {}
    for spec in specs:
        b, n = spec["batch"], spec["n"]
        data[(b, n)] = [harness.config_input(b, n, spec["seed"] + 42 * i, device=dev)
                        for i in range(harness.rotation_count(b, n))]
        print(f"# inputs {b}x{n} x{len(data[(b, n)])}", file=sys.stderr, flush=True)
    return data


def _gate(specs, data, mode, host_L=None):
    """Correctness of one form on every shape (untimed): the reference's gate
    and the FP64 residual ||LL^T - A||_F / ||A||_F.  host_L (a dict) receives
    a host copy of each factor for the forward-error comparison."""
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
        if host_L is not None:
            host_L.setdefault(key, {})[mode] = L.cpu()
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
            # Outside the timed region: let the previous shape's work drain
            # and the clocks settle (4096x32 follows 1x32768, ~50-100 ms of
            # full-chip load), then flush the caches.
            torch.cuda.synchronize()
            time.sleep(0.005)
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
    us, per_step = {}, {}
    for key, lst in ev.items():
        vals = [e0.elapsed_time(e1) * 1e3 / c for (e0, e1, c) in lst]
        us[key] = sum(vals) / len(vals)
        per_step[key] = vals
    return us, per_step


def _copy_gbps(inputs, dev):
    """A device copy of the same bytes as one factorisation of each input
    (read A, write an output of the same size), timed like the shapes: the
    practical ceiling of a kernel that reads and writes each byte once."""
    outs = [torch.empty_like(a) for a in inputs]
    ts = []
    for _ in range(6):
        torch.cuda.synchronize()
        harness.flush_caches(dev)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for a, o in zip(inputs, outs):
            o.copy_(a)
        e1.record()
        torch.cuda.synchronize()
        ts.append(e0.elapsed_time(e1) * 1e3 / len(inputs))
    ts.sort()
    return 2.0 * inputs[0].numel() * 4 / ts[len(ts) // 2] / 1e3


def _event_us(fn, reps=5):
    """Median microseconds of fn() over reps event-timed calls (cache flushed)."""
    dev = torch.device("cuda", torch.cuda.current_device())
    fn()
    torch.cuda.synchronize()
    ts = []
    for _ in range(reps):
        harness.flush_caches(dev)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        fn()
        e1.record()
        torch.cuda.synchronize()
        ts.append(e0.elapsed_time(e1) * 1e3)
    ts.sort()
    return ts[len(ts) // 2]


def extras(dev):
    """Ragged workload, graph replay and backward measurements (not part of
    the headline value)."""
    from gpumode_cholesky_amd.autograd import cholesky_ad
    from gpumode_cholesky_amd.graphs import GraphedCholesky
    from gpumode_cholesky_amd.ragged import cholesky_ragged

    out = {}
    # Ragged: 512 matrices, n log-uniform on [32, 2048], each G G^T/n + I.
    from gpumode_cholesky_amd.ragged_workload import workload
    mats = workload(dev)
    ns = np.array([a.shape[-1] for a in mats])
    gb = sum(a.numel() * 4 for a in mats) / 1e9
    rag = {"count": len(mats), "gb": round(gb, 3), "n_min": int(ns.min()), "n_max": int(ns.max())}
    for mode in ("plain", "mixed"):
        rag[f"one_call_us_{mode}"] = round(_event_us(lambda: cholesky_ragged(mats, mode=mode), 7), 1)
    rag["per_matrix_calls_us_plain"] = round(_event_us(lambda: [cholesky(a[None]) for a in mats], 3), 1)
    out["ragged"] = rag
    del mats
    torch.cuda.empty_cache()
    print("# extras: ragged done", file=sys.stderr, flush=True)

    # The fused route (one workgroup per matrix) on a shape it takes by
    # default, against the same shape forced onto the blocked path.
    from gpumode_cholesky_amd import _lib, autotune
    fu = {}
    for b, n in ((2000, 300),):
        a = harness.dense_input(b, n, 1.0, 2300, device=dev)
        row = {"route": int(_lib.load().chol_route(b, n, 0, None))}
        for mode in ("plain", "mixed"):
            row[f"fused_us_{mode}"] = round(_event_us(lambda: cholesky(a, mode=mode), 5), 1)
            try:
                autotune.set_table([[n, n, b, b, 256]])
                row[f"blocked_us_{mode}"] = round(_event_us(lambda: cholesky(a, mode=mode), 5), 1)
            finally:
                autotune.load_table()
        fu[f"{b}x{n}"] = row
        del a
    out["fused_route"] = fu
    print("# extras: fused route done", file=sys.stderr, flush=True)

    # Graph replay vs eager calls: the two smallest fixed shapes, and calls
    # that leave most CUs idle.  Each shape's rotated inputs (as many as the
    # timed protocol uses, min(50, 256 MiB / input)) in one graph of
    # sequential calls, and in one of independent branches.
    gr = {}
    for b, n in ((4096, 32), (1024, 64), (16, 128), (4, 256)):
        ins = [harness.config_input(b, n, 31 + i, device=dev) for i in range(harness.rotation_count(b, n))]
        g = GraphedCholesky(ins)
        gb = GraphedCholesky(ins, branches=True)
        eager = _event_us(lambda: [cholesky(a) for a in ins], 9) / len(ins)
        replay = _event_us(g.replay, 9) / len(ins)
        replay_b = _event_us(gb.replay, 9) / len(ins)
        gr[f"{b}x{n}"] = {"inputs": len(ins), "eager_us": round(eager, 2), "graph_us": round(replay, 2),
                          "graph_branches_us": round(replay_b, 2)}
        del g, gb, ins
    out["graph"] = gr
    print("# extras: graph done", file=sys.stderr, flush=True)

    # Forward + backward (spectral cond 1e2), L^-1 recomputed vs stored.
    bw = {}
    for b, n in ((256, 128), (8, 2048)):
        a = harness.spectral_input(b, n, 1e2, 55, device=dev)
        gL = torch.tril(torch.randn(b, n, n, device=dev, generator=torch.Generator(device=dev).manual_seed(1)))
        row = {}
        for store in (False, True):
            A = a.clone().requires_grad_(True)

            def step():
                A.grad = None
                L = cholesky_ad(A, store_inverse=store)
                L.backward(gL)

            torch.cuda.synchronize()
            torch.cuda.reset_peak_memory_stats(dev)
            base = torch.cuda.memory_allocated(dev)
            us = _event_us(step, 3)
            peak = (torch.cuda.max_memory_allocated(dev) - base) / 2**20
            row["stored" if store else "recomputed"] = {"fwd_bwd_us": round(us, 1), "peak_extra_mib": round(peak, 1)}
        bw[f"{b}x{n}"] = row
        del a, gL
    out["backward"] = bw
    print("# extras: backward done", file=sys.stderr, flush=True)

    # The fp8 form (E4M3 bulk updates from n = 16384 on) against the mixed
    # form on the two largest shapes: time per call (median of 3, caches
    # flushed), the reference's gate and the FP64 residual.
    f8 = {}
    for b, n in ((1, 16384), (1, 32768)):
        a = harness.config_input(b, n, 4700 + n, device=dev)
        row = {}
        for mode in ("mixed", "fp8"):
            row[f"{mode}_us"] = round(_event_us(lambda: cholesky(a, mode=mode), 3), 1)
            L = cholesky(a, mode=mode)
            ok, _ = harness.check_implementation(a, L)
            row[f"{mode}_gate"] = bool(ok)
            row[f"{mode}_fp64_resid"] = harness.fp64_residual(a, L)
            del L
            torch.cuda.empty_cache()
        f8[f"{b}x{n}"] = row
        del a
        torch.cuda.empty_cache()
    out["fp8_form"] = f8
    print("# extras: fp8 form done", file=sys.stderr, flush=True)
    return out


def paired(specs, dev):
    """A-B-B-A same-process comparison, plain (A) against mixed (B), per
    shape (scripts/_gpu_runner.py:2494-2588): median ratio plain/mixed with
    its bootstrap 95 % interval; one JSON line."""
    rows = []
    for spec in specs:
        b, n = spec["batch"], spec["n"]
        k, iters, repeats = harness.paired_budget(n)
        data = [harness.config_input(b, n, spec["seed"] + 42 * i, device=dev) for i in range(k)]
        r = harness.paired_compare(lambda d: cholesky(d, mode="plain"), lambda d: cholesky(d, mode="mixed"),
                                   data, iters, repeats, warmup=3 if n <= 2048 else 2)
        r.update({"batch": b, "n": n})
        rows.append(r)
        print(f"# paired {b}x{n}: plain/mixed {r['ratio_median']:.3f} "
              f"[{r['ratio_ci95'][0]:.3f}, {r['ratio_ci95'][1]:.3f}] plain {r['a_us']:.1f} us mixed {r['b_us']:.1f} us",
              file=sys.stderr, flush=True)
        del data
        torch.cuda.empty_cache()
    print(json.dumps({"mode": "paired", "a": "plain", "b": "mixed",
                      "geomean_ratio": _geo(r["ratio_median"] for r in rows), "rows": rows}))


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
    if args.copy_probe:
        src = torch.ones((args.copy_probe * 2**20 // 4,), device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)  # reads and writes exactly copy_probe MiB
        torch.cuda.synchronize()
        del src, dst
    if args.paired:
        paired(specs, dev)
        return
    modes = ["plain", "mixed"] if args.mode == "both" else [args.mode]
    data = _inputs(specs, dev)
    host_L = {} if not args.no_cpu_baseline else None
    resid = {m: _gate(specs, data, m, host_L) for m in modes}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    timed = {m: _time(specs, data, m, args.steps, args.warmup, dev) for m in modes}
    times = {m: timed[m][0] for m in modes}
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
                      "fp64_resid": resid[m][key],
                      "step_us": [round(v, 2) for v in timed[m][1][key]]}
        rows.append(row)
    geo = {m: _geo(times[m][k] for k in times[m]) for m in modes}
    ms_per_step = sum(times[head].values()) / 1e3

    # Headline roofline: the single-launch wave-register kernel on 4096x32
    # (one launch per call, so the per-call event time is its duration; both
    # forms run this same kernel, so the faster of the two per-call means).
    roof = None
    r32 = next((r for r in rows if (r["batch"], r["n"]) == (4096, 32)), None)
    if r32 is not None:
        g = max(r32[m]["GBps"] for m in modes)
        # traffic: HBM bytes per launch from the two rocprofv3 --pmc passes
        # (FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, scaled by a 256 MiB copy
        # probe), recorded in profiles/k_potrf_small_32_4096x32.md.
        roof = {"kernel": "k_potrf_small<32,vec> on 4096x32", "bound": "hbm",
                "achieved": g, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(g / HBM_PEAK_GBS, 4), "traffic": PMC_TRAFFIC_4096x32,
                "algorithmic_bytes": 8 * 4096 * 32 * 32,
                "copy_same_bytes_GBps": round(_copy_gbps(data[(4096, 32)], dev), 1)}
    cpu = fwd = None
    if not args.no_cpu_baseline:
        cpu, fwd = cpu_baseline(specs, data, host_L)
        for row in rows:
            for m in modes:
                row[m]["fwd_vs_torch_cpu"] = fwd[(row["batch"], row["n"])].get(m)
    out = {
        "metric": ("geomean us over 15 batch x n SPD shapes (mixed form; plain in geomean_us)" if head == "mixed"
                   else f"geomean us over 15 batch x n SPD shapes ({head} form)"),
        "value": round(geo[head], 2),
        "unit": "us",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"plain": "f32", "mixed": "f32+bf16x3", "fp8": "f32+bf16x3+e4m3"}[head],
        "data": "synthetic",
        "config": {"workload": "15-shape batched Cholesky grid (4096x32 ... 1x32768), FP32 in/out",
                   "form": head, "value_form": f"cholesky(A, mode='{head}')",
                   "drop_in_form": "custom_kernel(A) = cholesky(A, mode='plain')", "shapes": len(rows)},
        "roofline": roof,
        "geomean_us": {m: round(geo[m], 2) for m in modes},
        "max_fp64_resid": {m: max(resid[m].values()) for m in modes},
        "wall_s_timed": round(wall, 2),
        "per_shape": rows,
    }
    if len(modes) == 2:
        out["mixed_speedup_vs_plain"] = round(geo["plain"] / geo["mixed"], 4)
    if not args.no_extras:
        out["extras"] = extras(dev)
    if cpu is not None:
        out["cpu_baseline"] = cpu
    if args.detail:
        with open(args.detail, "w") as f:
            json.dump(out, f, indent=1)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
