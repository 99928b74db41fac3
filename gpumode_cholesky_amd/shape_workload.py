"""One (batch, n) shape as a runnable module, to time or profile a route:
K factorisations of the same seeded input, after one warmup call, in the
chosen form and route.  The input is dense G G^T/n + I (quick to generate
on the device; the kernels' time does not depend on the values), or with
--input config the benchmark's generator for the order (spectral cond 1e3
for 64 <= n <= 1024, built through a batched QR).

    python -m gpumode_cholesky_amd.shape_workload --batch 2000 --n 300 \
        [--mode mixed|plain|fp8|both] [--calls K] [--route default|fused|blocked|lds] [--nbo W] [--input dense|config]

--route forces the route through the tuning table for this shape only:
fused (the one-workgroup-per-matrix kernel), blocked (outer blocks of
--nbo columns, default 256), lds (the whole-matrix LDS kernel, n <= 256).
Prints the mean event-timed microseconds per call of each form.
"""
from __future__ import annotations

import argparse

import torch

from . import autotune, harness
from .cholesky import cholesky

_ROUTES = {"fused": -2, "lds": -1}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--batch", type=int, required=True)
    ap.add_argument("--n", type=int, required=True)
    ap.add_argument("--mode", default="both", choices=["plain", "mixed", "fp8", "both"])
    ap.add_argument("--calls", type=int, default=3)
    ap.add_argument("--route", default="default", choices=["default", "fused", "blocked", "lds"])
    ap.add_argument("--nbo", type=int, default=256)
    ap.add_argument("--input", default="dense", choices=["dense", "config"])
    args = ap.parse_args()
    dev = torch.device("cuda", 0)
    b, n = args.batch, args.n
    if args.input == "config":
        a = harness.config_input(b, n, 4242 + n, device=dev)
    else:
        a = harness.dense_input(b, n, 1.0, 4242 + n, device=dev)
    if args.route != "default":
        nbo = args.nbo if args.route == "blocked" else _ROUTES[args.route]
        autotune.set_table([[n, n, b, b, nbo]])
    modes = ["plain", "mixed"] if args.mode == "both" else [args.mode]
    try:
        for mode in modes:
            cholesky(a, mode=mode)  # warmup
            torch.cuda.synchronize()
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(args.calls):
                cholesky(a, mode=mode)
            e1.record()
            torch.cuda.synchronize()
            us = e0.elapsed_time(e1) * 1e3 / max(args.calls, 1)
            print(f"{b}x{n} {mode} route={args.route}: {us:.1f} us per call ({args.calls} calls)", flush=True)
    finally:
        autotune.load_table()


if __name__ == "__main__":
    main()
