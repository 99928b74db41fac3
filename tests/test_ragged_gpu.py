"""GPU parity of the ragged path (chol_potrf_ragged_f32).

Each matrix of a mixed-order list must match the FP64 oracle to the same
FP32 tolerance as the dense path (reconstruction <= 8 sqrt(n) eps32), be
bit-identical to the dense path where both run the same kernel arithmetic
(n <= 64), and report info per matrix in the caller's order.  Orders
65..512 run the fused kernel by default and the lock-step blocked path when
the tuning table routes them elsewhere (LOCKSTEP); both are tested.
"""
import math

import numpy as np
import pytest
import torch

from gpumode_cholesky_amd import autotune, cholesky, harness
from gpumode_cholesky_amd.ragged import cholesky_ragged

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda", 0)
EPS32 = float(np.finfo(np.float32).eps)

SIZES = [1, 2, 16, 17, 31, 32, 33, 48, 64, 65, 100, 128, 129, 200, 256, 257, 300, 512, 700, 1024, 5, 64, 32]
# Every order above 64 on the lock-step path (outer blocks of 256).
LOCKSTEP = [[65, 32768, 1, 1 << 30, 256]]


def _with_table(table, fn):
    try:
        if table is not None:
            autotune.set_table(table)
        return fn()
    finally:
        autotune.load_table()


@pytest.mark.parametrize("route", ["default", "lockstep"])
def test_ragged_mixed_orders_against_oracle(oracle_mod, route):
    mats = [harness.generate_input(1, n, 2, 700 + i, device="cpu")[0] for i, n in enumerate(SIZES)]
    outs, info = _with_table(LOCKSTEP if route == "lockstep" else None,
                             lambda: cholesky_ragged([m.to(DEV) for m in mats], return_info=True))
    assert (info.cpu() == 0).all()
    for i, (a, L) in enumerate(zip(mats, outs)):
        n = a.shape[0]
        Lc = L.cpu().numpy()
        assert (np.triu(Lc, 1) == 0).all(), f"matrix {i} n={n}: non-zero above the diagonal"
        fro = oracle_mod.residual(a.numpy()[None], Lc[None])[0][0]
        print(f"[ragged] #{i} n={n}: resid={fro:.3e}", flush=True)
        assert fro <= 8 * math.sqrt(n) * EPS32, f"matrix {i} n={n}: residual {fro:.3e}"
        if n <= 64:
            dense = cholesky(a.to(DEV)[None])[0]
            assert torch.equal(dense, L), f"matrix {i} n={n}: differs from the dense path"


@pytest.mark.parametrize("route", ["default", "lockstep"])
@pytest.mark.parametrize("sizes", [[65, 100, 128, 77, 128], [129, 200, 256, 65, 100, 150, 255, 33]])
def test_ragged_mid_orders_only(oracle_mod, sizes, route):
    # Mid orders alone: the fused kernel (default; no lock-step class, so no
    # side stream), or the lock-step path, where with no larger matrix the
    # block that ends the largest is every matrix's last (no inverse).
    mats = [harness.generate_input(1, n, 3, 900 + i, device="cpu")[0] for i, n in enumerate(sizes)]
    outs, info = _with_table(LOCKSTEP if route == "lockstep" else None,
                             lambda: cholesky_ragged([m.to(DEV) for m in mats], return_info=True))
    assert (info.cpu() == 0).all()
    for i, (a, L) in enumerate(zip(mats, outs)):
        n = a.shape[0]
        Lc = L.cpu().numpy()
        assert (np.triu(Lc, 1) == 0).all(), f"matrix {i} n={n}: non-zero above the diagonal"
        fro = oracle_mod.residual(a.numpy()[None], Lc[None])[0][0]
        assert fro <= 8 * math.sqrt(n) * EPS32, f"matrix {i} n={n}: residual {fro:.3e}"


def test_ragged_info_in_caller_order():
    sizes = [300, 40, 7, 150, 64]
    mats = [harness.generate_input(1, n, 2, 50 + n, device="cpu")[0] for n in sizes]
    bad = {1: 10, 3: 77, 0: 200}  # matrix index -> pivot made negative (0-based)
    for i, k in bad.items():
        mats[i][k, k] = -1.0
    outs, info = cholesky_ragged([m.to(DEV) for m in mats], return_info=True)
    info = info.cpu().tolist()
    for i in range(len(sizes)):
        assert info[i] == (bad[i] + 1 if i in bad else 0), (i, info)
        if i not in bad:
            assert torch.isfinite(outs[i]).all()


def test_ragged_empty_and_single():
    assert cholesky_ragged([]) == []
    a = harness.generate_input(1, 50, 2, 1, device=DEV)[0]
    (L,) = cholesky_ragged([a])
    assert torch.equal(L, cholesky(a[None])[0])


def ragged_workload(count=512, lo=32, hi=2048, seed=4096):
    """512 SPD matrices, n log-uniform on [32, 2048], each G G^T/n + I."""
    rng = np.random.default_rng(seed)
    ns = np.exp(rng.uniform(math.log(lo), math.log(hi), count)).astype(int)
    return [harness.dense_input(1, int(n), 1.0, seed + i, device=DEV)[0] for i, n in enumerate(ns)]


@pytest.mark.parametrize("mode", ["plain", "mixed"])
def test_ragged_workload_gate(mode):
    mats = ragged_workload()
    outs, info = cholesky_ragged(mats, mode=mode, return_info=True)
    torch.cuda.synchronize()
    assert (info == 0).all()
    worst = 0.0
    for a, L in zip(mats, outs):
        ok, msg = harness.check_implementation(a[None], L[None])
        assert ok, f"n={a.shape[0]}: {msg}"
        worst = max(worst, harness.fp64_residual(a[None], L[None]))
    total = sum(a.numel() * 4 for a in mats) / 1e9
    print(f"[ragged-workload] {mode}: {len(mats)} matrices, {total:.2f} GB, worst fp64 resid {worst:.3e}", flush=True)
    assert worst <= (1e-5 if mode == "mixed" else 8 * math.sqrt(2048) * EPS32)


@pytest.mark.parametrize("route", ["default", "lockstep"])
def test_ragged_more_than_65535_large_matrices(route):
    # 65600 matrices of orders 65..72: on the lock-step path (whose kernels
    # put the matrix on grid.y) the call splits into chunks of at most
    # 65535; the fused kernel takes them in one launch (matrix on grid.x).
    # Sampled factors on both sides of the split, and info.
    count = 65600
    ns = [65 + (i % 8) for i in range(count)]
    base = {n: harness.generate_input(1, n, 2, 880 + n, device=DEV)[0] for n in set(ns)}
    mats = [base[n] for n in ns]
    outs, info = _with_table(LOCKSTEP if route == "lockstep" else None,
                             lambda: cholesky_ragged(mats, return_info=True))
    torch.cuda.synchronize()
    assert (info == 0).all().item()
    for i in (0, 7, 65534, 65535, 65536, count - 1):
        n = ns[i]
        ref = cholesky(base[n][None])[0]
        d = ((outs[i] - ref).abs().max() / ref.abs().max()).item()
        r = harness.fp64_residual(base[n][None], outs[i].contiguous()[None])
        print(f"[ragged>65535] matrix {i} n={n}: vs dense {d:.2e} resid {r:.2e}", flush=True)
        assert d < 1e-5 and r <= 8 * math.sqrt(n) * EPS32, f"matrix {i} (n={n})"
