"""Autotuner for the route / blocking table of the C library.

The library picks, per (n, batch), either the whole-matrix LDS kernel
(n <= 256) or the blocked path with an outer block width ``nbo`` (columns per
trailing SYRK).  This module times the candidates on the GPU, keeps the
fastest per shape, and writes a JSON table that ``load_table`` (called when
the library is first loaded) hands to ``chol_set_tuning``.  It replaces the
reference's hand-maintained per-shape tables (submission.py:2263-2281,
:3262-3281, :5364-5371).

    python -m gpumode_cholesky_amd.autotune [--out PATH] [--shapes 4x1024,...]
"""
from __future__ import annotations

import argparse
import json
import os

import numpy as np
import torch

from . import _lib

DEFAULT_TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json")
NBO_CANDIDATES = (128, 256, 512, 1024, 2048)
FUSED = -2  # table value of the fused route
FUSED_MAX_N = 1024  # largest order the tuner times on the fused route
kNB_MIN = 128  # the smallest outer width (one diagonal block)


def set_table(rows) -> None:
    """Install rows [n_lo, n_hi, batch_lo, batch_hi, mode, nbo] (mode 0 plain,
    1 mixed, -1 both; a 5-element row applies to both forms).  First match
    wins."""
    rows = [list(r) if len(r) == 6 else list(r[:4]) + [-1, r[4]] for r in rows]
    arr = np.ascontiguousarray(np.asarray(rows, dtype=np.int64).reshape(-1, 6))
    _lib.check(_lib.load().chol_set_tuning(arr.ctypes.data if len(arr) else None, len(arr)),
               "chol_set_tuning")


def load_table(path: str = DEFAULT_TABLE) -> list:
    """Install the table stored at ``path`` (no-op when the file is absent)."""
    if not os.path.exists(path):
        return []
    with open(path) as f:
        rows = json.load(f)["rows"]
    set_table(rows)
    return rows


def candidates(n: int):
    """Table values to time for order n: -1 (whole-matrix LDS kernel,
    n <= 256), -2 (fused one-workgroup-per-matrix kernel, 64 < n <= 1024:
    above that one matrix on one workgroup takes milliseconds to seconds)
    and the blocked path's outer widths up to n."""
    out = [-1] if n <= 256 else []
    if 64 < n <= FUSED_MAX_N:
        out.append(FUSED)
    if n > 64:
        out += [c for c in NBO_CANDIDATES if c <= max(kNB_MIN, n)]
    return out


def time_once(a: torch.Tensor, mode: str) -> float:
    """Microseconds of one event-timed call (caches flushed first)."""
    from .cholesky import cholesky
    from .harness import flush_caches

    flush_caches(a.device)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    cholesky(a, mode=mode)
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) * 1e3


def time_candidates(a: torch.Tensor, mode: str, cands, reps: int = 5) -> dict:
    """Median microseconds per call of every candidate table value for the
    shape of ``a``.  Each is warmed up first; then the timed calls go
    round-robin over the candidates, the order reversed every round, so a
    drift of the GPU's clock or memory state over the run (the first
    calls of a process ran a few per cent slower) weighs on all of them
    alike instead of on whichever is timed first."""
    from .cholesky import cholesky

    batch, n = a.shape[0], a.shape[-1]
    for nbo in cands:
        set_table([[n, n, batch, batch, nbo]])
        cholesky(a, mode=mode)
    torch.cuda.synchronize()
    ts = {nbo: [] for nbo in cands}
    for r in range(reps):
        for nbo in (cands if r % 2 == 0 else cands[::-1]):
            set_table([[n, n, batch, batch, nbo]])
            ts[nbo].append(time_once(a, mode))
    return {nbo: float(np.median(v)) for nbo, v in ts.items()}


def tune(shapes, mode: str = "plain", reps: int = 5, log=print) -> list:
    """Time every candidate for each (batch, n) in one form; return the
    winning rows (tagged with that form)."""
    from .harness import config_input

    mcode = 1 if mode == "mixed" else 0
    rows = []
    dev = torch.device("cuda", torch.cuda.current_device())
    for batch, n in shapes:
        if n <= 64:
            continue
        a = config_input(batch, n, 1234 + n, device=dev)
        times = time_candidates(a, mode, candidates(n), reps)
        for nbo, us in times.items():
            log(f"tune {batch}x{n} {mode} nbo={nbo}: {us:.1f} us")
        best = min(times.items(), key=lambda kv: kv[1])
        rows.append([n, n, batch, batch, mcode, best[0]])
        log(f"tune {batch}x{n} {mode}: best nbo={best[0]} ({best[1]:.1f} us)")
        del a
        torch.cuda.empty_cache()
    set_table([])
    return rows


def main():
    from .harness import BENCH_SPECS

    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=DEFAULT_TABLE)
    ap.add_argument("--shapes", default="")
    ap.add_argument("--mode", default="plain", choices=["plain", "mixed"])
    ap.add_argument("--reps", type=int, default=5)
    args = ap.parse_args()
    if args.shapes:
        shapes = [tuple(int(x) for x in s.split("x")) for s in args.shapes.split(",") if s]
    else:
        shapes = [(s["batch"], s["n"]) for s in BENCH_SPECS]
    rows = tune(shapes, args.mode, args.reps, log=lambda m: print(m, flush=True))
    # keep the other form's rows of an existing table
    mcode = 1 if args.mode == "mixed" else 0
    keep = []
    if os.path.exists(args.out):
        with open(args.out) as f:
            keep = [r for r in json.load(f)["rows"] if len(r) == 6 and r[4] != mcode]
    with open(args.out, "w") as f:
        json.dump({"columns": ["n_lo", "n_hi", "batch_lo", "batch_hi", "mode", "nbo"], "rows": keep + rows}, f)
    print(f"wrote {len(rows)} {args.mode} rows to {args.out} ({len(keep)} kept)")


if __name__ == "__main__":
    main()
