"""The caller side of ``custom_kernel``: input generators, the correctness
gate and the timing protocol, restated from the reference's harness so the
engine can be tested and benchmarked the way the reference is.

* ``generate_input`` — the six seeded SPD families of
  reference/reference.py:37-126 (same torch op sequence, so the same bits on
  the same device/generator).
* ``config_input`` — the synthetic workloads the benchmark is quoted on:
  dense ``G G^T/n + d I`` and spectral ``Q diag(lambda) Q^T`` matrices.
* ``check_implementation`` — the gate of reference/reference.py:137-188.
* ``BENCH_SPECS`` — the 15 ranked shapes and seeds
  (scripts/_gpu_runner.py:112-128).
* ``benchmark`` — the eval.py protocol (reference/eval.py:132-200): rotate
  ``min(50, 256 MiB / input)`` inputs, flush the caches before each repeat,
  time the per-input loop with events, report the mean per call.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass

import torch

# reference/reference.py:6-7
RECON_RTOL_FACTOR = 20.0
TRIANGULAR_RTOL_FACTOR = 8.0

# scripts/_gpu_runner.py:112-128 (the 15-shape ranked grid; cond 2 = dense).
BENCH_SPECS = [
    {"batch": 4096, "n": 32, "cond": 2, "seed": 41032},
    {"batch": 1024, "n": 64, "cond": 2, "seed": 41064},
    {"batch": 256, "n": 128, "cond": 2, "seed": 41128},
    {"batch": 64, "n": 256, "cond": 2, "seed": 41256},
    {"batch": 16, "n": 512, "cond": 2, "seed": 41512},
    {"batch": 640, "n": 512, "cond": 2, "seed": 510512},
    {"batch": 4, "n": 1024, "cond": 2, "seed": 42024},
    {"batch": 60, "n": 1024, "cond": 2, "seed": 511024},
    {"batch": 2, "n": 2048, "cond": 2, "seed": 44048},
    {"batch": 8, "n": 2048, "cond": 2, "seed": 512048},
    {"batch": 1, "n": 4096, "cond": 2, "seed": 48096},
    {"batch": 2, "n": 4096, "cond": 2, "seed": 514096},
    {"batch": 1, "n": 8192, "cond": 2, "seed": 48192},
    {"batch": 1, "n": 16384, "cond": 2, "seed": 48284},
    {"batch": 1, "n": 32768, "cond": 2, "seed": 48368},
]

# scripts/_gpu_runner.py:2966-2973 — family and cond of the regression grid.
FAMILIES = [
    ("dense", 2),
    ("spectrum", 5),
    ("diagonal", 5),
    ("lowrank", 4),
    ("rowscale", 4),
    ("tridiagonal", 1),
]


def _default_device() -> torch.device:
    return torch.device("cuda" if torch.cuda.is_available() else "cpu")


def _symmetrize(a: torch.Tensor) -> torch.Tensor:
    # reference/reference.py:10-11
    return 0.5 * (a + a.transpose(-1, -2))


def _random_orthogonal(batch: int, n: int, gen: torch.Generator,
                       device: torch.device) -> torch.Tensor:
    # reference/reference.py:14-27: Q of a Gaussian's QR, columns sign-fixed
    # so diag(R) >= 0.
    x = torch.randn((batch, n, n), device=device, dtype=torch.float32, generator=gen)
    q, r = torch.linalg.qr(x)
    signs = torch.where(torch.diagonal(r, dim1=-2, dim2=-1) >= 0,
                        torch.ones((), device=device), -torch.ones((), device=device))
    return q * signs.unsqueeze(-2)


def _dense_spd(a: torch.Tensor, damping: float) -> torch.Tensor:
    # reference/reference.py:30-34
    n = a.shape[-1]
    out = (a @ a.transpose(-1, -2)) / float(max(n, 1))
    out.diagonal(dim1=-2, dim2=-1).add_(damping)
    return _symmetrize(out).contiguous()


def generate_input(batch: int, n: int, cond: int, seed: int, case: str = "dense",
                   device: torch.device | str | None = None) -> torch.Tensor:
    """Seeded SPD batch of one family (reference/reference.py:37-126).

    ``device`` defaults to the reference's choice (cuda when available).
    """
    assert batch > 0, "batch must be positive"
    assert n > 0, "n must be positive"
    assert cond >= 0, "cond must be non-negative"
    device = torch.device(device) if device is not None else _default_device()
    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    case = case.lower()
    if case == "dense":  # :52-57
        a = torch.randn((batch, n, n), device=device, dtype=torch.float32, generator=gen)
        return _dense_spd(a, 10.0 ** -max(cond, 2))
    if case == "diagonal":  # :59-67
        d = torch.logspace(0.0, -float(max(cond, 1)), n, device=device, dtype=torch.float32)
        return torch.diag_embed(d.expand(batch, n)).contiguous()
    if case == "spectrum":  # :69-80
        values = torch.logspace(0.0, -float(max(cond, 1)), n, device=device,
                                dtype=torch.float32).expand(batch, n)
        q = _random_orthogonal(batch, n, gen, device)
        return _symmetrize((q * values.unsqueeze(-2)) @ q.transpose(-1, -2)).contiguous()
    if case == "lowrank":  # :82-92
        rank = max(1, min(n // 8, 64))
        f = torch.randn((batch, n, rank), device=device, dtype=torch.float32, generator=gen)
        out = (f @ f.transpose(-1, -2)) / float(rank)
        out.diagonal(dim1=-2, dim2=-1).add_(10.0 ** -max(cond, 3))
        return _symmetrize(out).contiguous()
    if case == "rowscale":  # :94-108
        a = torch.randn((batch, n, n), device=device, dtype=torch.float32, generator=gen)
        base = _dense_spd(a, 1.0e-3)
        scales = torch.logspace(0.0, -0.5 * float(max(cond, 2)), n, device=device,
                                dtype=torch.float32)
        out = scales.reshape(1, n, 1) * base * scales.reshape(1, 1, n)
        out.diagonal(dim1=-2, dim2=-1).add_(1.0e-6)
        return _symmetrize(out).contiguous()
    if case == "tridiagonal":  # :110-124
        d = torch.empty((batch, n), device=device, dtype=torch.float32).uniform_(
            1.5, 2.5, generator=gen)
        off = torch.empty((batch, max(n - 1, 0)), device=device,
                          dtype=torch.float32).uniform_(-0.25, 0.25, generator=gen)
        out = torch.diag_embed(d)
        if n > 1:
            out = out + torch.diag_embed(off, offset=1) + torch.diag_embed(off, offset=-1)
        return out.contiguous()
    raise ValueError(f"unknown Cholesky test case: {case}")


def dense_input(batch: int, n: int, damping: float, seed: int,
                device: torch.device | str | None = None) -> torch.Tensor:
    """``G G^T / n + damping * I`` with ``G ~ N(0,1)`` from ``seed``, symmetrised
    (the dense workloads of the benchmark: damping 1.0 at n = 32, 0.1 at
    n >= 2048)."""
    device = torch.device(device) if device is not None else _default_device()
    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    g = torch.randn((batch, n, n), device=device, dtype=torch.float32, generator=gen)
    return _dense_spd(g, damping)


def spectral_input(batch: int, n: int, cond: float, seed: int,
                   device: torch.device | str | None = None) -> torch.Tensor:
    """``Q diag(lambda) Q^T`` with ``lambda = logspace(0, -log10(cond), n)`` and
    ``Q`` from the QR of a seeded Gaussian (sign-fixed), symmetrised."""
    device = torch.device(device) if device is not None else _default_device()
    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    values = torch.logspace(0.0, -math.log10(cond), n, device=device,
                            dtype=torch.float32).expand(batch, n)
    q = _random_orthogonal(batch, n, gen, device)
    return _symmetrize((q * values.unsqueeze(-2)) @ q.transpose(-1, -2)).contiguous()


def config_input(batch: int, n: int, seed: int, device=None) -> torch.Tensor:
    """The benchmark workload at one of the 15 shapes: n = 32 dense with
    damping 1.0; 64 <= n <= 1024 spectral with cond 1e3; n >= 2048 dense with
    damping 0.1."""
    if n <= 32:
        return dense_input(batch, n, 1.0, seed, device)
    if n <= 1024:
        return spectral_input(batch, n, 1.0e3, seed, device)
    return dense_input(batch, n, 0.1, seed, device)


def _l1(m: torch.Tensor) -> torch.Tensor:
    # reference/reference.py:133-134
    return torch.linalg.matrix_norm(m, ord=1, dim=(-2, -1))


def check_implementation(data: torch.Tensor, output: torch.Tensor):
    """The reference's gate (reference/reference.py:137-188), on device.

    Returns ``(ok, message)``.  The reconstruction product runs in FP32 with
    TF32-style reduced precision off.
    """
    if not isinstance(output, torch.Tensor):
        return False, "output must be a torch.Tensor"
    if output.shape != data.shape:
        return False, f"output shape must be {tuple(data.shape)}, got {tuple(output.shape)}"
    if output.dtype != torch.float32:
        return False, f"output dtype must be torch.float32, got {output.dtype}"
    if output.device != data.device:
        return False, f"output must be on {data.device}, got {output.device}"
    if not torch.isfinite(output).all().item():
        return False, "output contains NaN or Inf"
    batch, n, _ = data.shape
    eps = torch.finfo(torch.float32).eps
    scale = _l1(data).clamp_min(torch.finfo(torch.float32).tiny)
    upper = torch.triu(output, diagonal=1)
    tri = _l1(upper)
    if torch.any(tri > TRIANGULAR_RTOL_FACTOR * max(n, 1) * eps * scale).item():
        worst = (tri / scale).amax().item()
        return False, f"output is not lower triangular enough: relative_residual={worst:.3g}"
    diag = torch.diagonal(output, dim1=-2, dim2=-1)
    if torch.any(diag <= 0).item():
        return False, "output diagonal must be strictly positive"
    old = torch.backends.cuda.matmul.allow_tf32
    try:
        torch.backends.cuda.matmul.allow_tf32 = False
        recon = output @ output.transpose(-1, -2)
    finally:
        torch.backends.cuda.matmul.allow_tf32 = old
    res = _l1(recon - data)
    if torch.any(res > RECON_RTOL_FACTOR * max(n, 1) * eps * scale).item():
        worst = (res / scale).amax().item()
        return False, f"L @ L.T does not reconstruct A: relative_residual={worst:.3g}"
    scaled = (res / (eps * max(n, 1) * scale)).amax().item()
    return True, f"scaled_reconstruction_residual={scaled:.3g}; batch={batch}; n={n}"


def fp64_residual(data: torch.Tensor, output: torch.Tensor, chunk: int = 1):
    """``max_b ||L L^T - A||_F / ||A||_F`` formed in FP64 on the tensors'
    device (chunked over the batch to bound memory)."""
    worst = 0.0
    for s in range(0, data.shape[0], chunk):
        a = data[s:s + chunk].double()
        l = output[s:s + chunk].double()
        r = torch.linalg.matrix_norm(l @ l.transpose(-1, -2) - a) / torch.linalg.matrix_norm(a)
        worst = max(worst, r.max().item())
        del a, l, r
    return worst


_FLUSH = {}


def flush_caches(device: torch.device, mib: int = 512) -> None:
    """Overwrite ``mib`` MiB so neither L2 nor the 256 MiB Infinity Cache holds
    the inputs (reference/utils.py:169-176 writes 256 MiB for a B200's L2)."""
    buf = _FLUSH.get(device)
    if buf is None:
        buf = torch.empty((mib * 1024 * 1024 // 8,), dtype=torch.int64, device=device)
        _FLUSH[device] = buf
    buf.fill_(42)


@dataclass
class Stats:
    runs: int
    mean: float
    std: float
    err: float
    best: float
    worst: float


def calculate_stats(durations):
    # reference/eval.py:86-99
    runs = len(durations)
    mean = sum(durations) / runs
    var = sum((d - mean) ** 2 for d in durations)
    std = math.sqrt(var / (runs - 1)) if runs > 1 else 0.0
    return Stats(runs, mean, std, std / math.sqrt(runs), min(durations), max(durations))


def rotation_count(batch: int, n: int) -> int:
    # reference/eval.py:132-141
    return max(1, min(50, (256 * 1024 * 1024) // (batch * n * n * 4)))


def benchmark(fn, data_list, max_repeats: int = 200, max_time_s: float = 10.0,
              min_repeats: int = 3):
    """eval.py protocol (reference/eval.py:155-200), times in microseconds per call."""
    dev = data_list[0].device
    durations = []
    t0 = time.perf_counter()
    for it in range(max_repeats):
        torch.cuda.synchronize(dev)
        flush_caches(dev)
        start = torch.cuda.Event(enable_timing=True)
        end = torch.cuda.Event(enable_timing=True)
        start.record()
        for d in data_list:
            fn(d)
        end.record()
        torch.cuda.synchronize(dev)
        durations.append(start.elapsed_time(end) * 1e3 / len(data_list))
        elapsed = time.perf_counter() - t0
        if it + 1 >= min_repeats and elapsed > 0.1:
            st = calculate_stats(durations)
            if st.err / st.mean < 0.001 or st.mean * 1e-6 * st.runs > max_time_s or elapsed > 120:
                break
    return calculate_stats(durations)


def _median(values):
    s = sorted(values)
    m = len(s) // 2
    return s[m] if len(s) % 2 else 0.5 * (s[m - 1] + s[m])


def _bootstrap_ci(values, iters: int = 4000, seed: int = 20350719, alpha: float = 0.05):
    """Percentile bootstrap interval of the median (scripts/_gpu_runner.py:2438-2452)."""
    import random
    if len(values) < 2:
        return values[0], values[0]
    rng = random.Random(seed)
    reps = sorted(_median([values[rng.randrange(len(values))] for _ in values]) for _ in range(iters))
    return reps[int(alpha / 2 * iters)], reps[min(iters - 1, int((1 - alpha / 2) * iters))]


def paired_budget(n: int):
    """(rotating inputs, timed regions per block, repeats) by order
    (scripts/_gpu_runner.py:2461-2473)."""
    if n <= 128:
        return 2, 20, 9
    if n <= 512:
        return 2, 15, 9
    if n <= 2048:
        return 2, 10, 7
    if n <= 8192:
        return 2, 5, 7
    return 1, 3, 5


def timed_block(fn, data_list, iters: int) -> float:
    """Median us per call over ``iters`` cache-flushed, event-timed regions
    (scripts/_gpu_runner.py:2476-2491)."""
    dev = data_list[0].device
    out = []
    for _ in range(iters):
        flush_caches(dev)
        torch.cuda.synchronize(dev)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for d in data_list:
            fn(d)
        e1.record()
        torch.cuda.synchronize(dev)
        out.append(e0.elapsed_time(e1) * 1e3 / len(data_list))
    return _median(out)


def paired_compare(fn_a, fn_b, data_list, iters: int, repeats: int, warmup: int = 2) -> dict:
    """Same-process A-B-B-A comparison of two callables on the same inputs
    (scripts/_gpu_runner.py:2494-2588): both are warmed first, then each
    repeat times A, B, B, A blocks; the first repeat is discarded.  Reports
    the median of the per-repeat ratios mean(A)/mean(B) (> 1: B faster) with
    a bootstrap 95 % interval, and the A-vs-A spread inside a repeat (the
    protocol's own noise floor)."""
    for _ in range(warmup):
        for d in data_list:
            fn_a(d)
        for d in data_list:
            fn_b(d)
    torch.cuda.synchronize(data_list[0].device)
    ratios, a_us, b_us, spread = [], [], [], []
    for rep in range(repeats + 1):
        a1 = timed_block(fn_a, data_list, iters)
        b1 = timed_block(fn_b, data_list, iters)
        b2 = timed_block(fn_b, data_list, iters)
        a2 = timed_block(fn_a, data_list, iters)
        if rep == 0:
            continue
        am, bm = 0.5 * (a1 + a2), 0.5 * (b1 + b2)
        ratios.append(am / bm)
        a_us.append(am)
        b_us.append(bm)
        spread.append(abs(a1 - a2) / am)
    lo, hi = _bootstrap_ci(ratios)
    return {"ratio_median": _median(ratios), "ratio_ci95": [lo, hi], "a_us": _median(a_us),
            "b_us": _median(b_us), "order_spread_median": _median(spread), "repeats": len(ratios),
            "iters_per_block": iters, "rotating_inputs": len(data_list)}


# ---------------------------------------------------------------- stress ----
# Adversarial inputs of the reference's stress grids, cheap at every n (no
# QR): scripts/_gpu_runner.py:3051-3069 and :3374-3410.

def adversarial_diagonal(batch: int, n: int, exponent: int, device=None):
    """Diagonal SPD with eigenvalues logspace(0, -exponent, n); for exponent
    10 the largest and smallest eigenvalues are interleaved (near-singular,
    tiny diagonal and mixed dynamic range at once).  Returns (A, case name)
    (scripts/_gpu_runner.py:3051-3069)."""
    device = torch.device(device) if device is not None else _default_device()
    values = torch.logspace(0.0, -float(exponent), n, device=device, dtype=torch.float32)
    case = f"high_kappa_1e{exponent}"
    if exponent == 10:
        order = torch.empty(n, device=device, dtype=torch.long)
        half = (n + 1) // 2  # n // 2 for the even orders the reference uses
        order[0::2] = torch.arange(half, device=device)
        order[1::2] = torch.arange(n - 1, half - 1, -1, device=device)
        values = values[order]
        case = "mixed_dynamic_near_singular_1e10"
    return torch.diag_embed(values.unsqueeze(0).expand(batch, -1)).contiguous(), case


def mixed_dynamic_spd(batch: int, n: int, cond: int, seed: int, device=None):
    """D T D with T tridiagonal (diagonal U(1.75, 2.25), off-diagonal
    U(-0.08, 0.08)) and D = diag(logspace(0, -cond/2, n)): SPD with ~10^cond
    dynamic range (scripts/_gpu_runner.py:3374-3392)."""
    device = torch.device(device) if device is not None else _default_device()
    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    d = torch.empty((batch, n), device=device, dtype=torch.float32).uniform_(1.75, 2.25, generator=gen)
    off = torch.empty((batch, max(n - 1, 0)), device=device, dtype=torch.float32).uniform_(-0.08, 0.08, generator=gen)
    base = torch.diag_embed(d)
    if n > 1:
        base = base + torch.diag_embed(off, offset=1) + torch.diag_embed(off, offset=-1)
    scales = torch.logspace(0.0, -0.5 * float(cond), n, device=device, dtype=torch.float32)
    return (scales.reshape(1, n, 1) * base * scales.reshape(1, 1, n)).contiguous()


def near_singular_banded(batch: int, n: int, cond: int, device=None):
    """Diagonal logspace(0, -cond, n) with every other neighbouring pair
    coupled by 0.125 sqrt(d_i d_{i+1}): SPD 2x2 blocks down to 10^-cond
    (scripts/_gpu_runner.py:3395-3410)."""
    device = torch.device(device) if device is not None else _default_device()
    values = torch.logspace(0.0, -float(cond), n, device=device, dtype=torch.float32)
    out = torch.diag_embed(values.expand(batch, n)).contiguous()
    if n > 1:
        pair = torch.arange(n - 1, device=device)
        coupled = (pair % 2) == 0
        off = torch.zeros(n - 1, device=device, dtype=torch.float32)
        off[coupled] = 0.125 * torch.sqrt(values[:-1][coupled] * values[1:][coupled])
        out = out + torch.diag_embed(off.expand(batch, n - 1), offset=1)
        out = out + torch.diag_embed(off.expand(batch, n - 1), offset=-1)
    return out.contiguous()


def stress_cases(batch: int, n: int, seed: int, device=None):
    """The adversarial set of one shape, as (name, A) pairs: for cond 6, 8
    and 10 the tiny-diagonal family, the near-singular banded and the D T D
    inputs (scripts/_gpu_runner.py:3454-3467), and the high-kappa diagonals
    1e6 / 1e8 / interleaved 1e10 (:3051-3069)."""
    for cond in (6, 8, 10):
        yield f"tiny_diagonal_1e{cond}", generate_input(batch, n, cond, seed + 7100 + cond, "diagonal", device)
        yield f"near_singular_banded_1e{cond}", near_singular_banded(batch, n, cond, device)
        yield f"mixed_dynamic_1e{cond}", mixed_dynamic_spd(batch, n, cond, seed + 9100 + cond, device)
    for e in (6, 8, 10):
        a, name = adversarial_diagonal(batch, n, e, device)
        yield name, a
