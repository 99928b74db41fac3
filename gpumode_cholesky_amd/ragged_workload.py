"""The ragged benchmark workload as a runnable module: 512 SPD matrices
(G G^T / n + I) with orders drawn log-uniformly from [32, 2048] (seed 4096),
factored in one chol_potrf_ragged_f32 call per form, after a warmup call.
Used by bench.py's ragged extra and to profile that call:

    python -m gpumode_cholesky_amd.ragged_workload [--mode mixed|plain|both] [--calls K]
"""
from __future__ import annotations

import argparse
import math

import numpy as np
import torch

from . import harness
from .ragged import cholesky_ragged


def workload(dev, count: int = 512, seed: int = 4096):
    rng = np.random.default_rng(seed)
    ns = np.exp(rng.uniform(math.log(32), math.log(2048), count)).astype(int)
    return [harness.dense_input(1, int(n), 1.0, seed + i, device=dev)[0] for i, n in enumerate(ns)]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--mode", default="both", choices=["plain", "mixed", "both"])
    ap.add_argument("--calls", type=int, default=1)
    args = ap.parse_args()
    dev = torch.device("cuda", 0)
    mats = workload(dev)
    modes = ["plain", "mixed"] if args.mode == "both" else [args.mode]
    for mode in modes:
        cholesky_ragged(mats, mode=mode)  # warmup
        torch.cuda.synchronize()
        for _ in range(args.calls):
            cholesky_ragged(mats, mode=mode)
        torch.cuda.synchronize()
        print(f"ragged {len(mats)} matrices, {mode}: {args.calls} call(s) done", flush=True)


if __name__ == "__main__":
    main()
