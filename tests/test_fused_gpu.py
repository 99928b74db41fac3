"""GPU parity of the fused route (k_potrf_fused: one workgroup factors a
whole matrix, every step's diagonal block, panel and trailing update in one
launch), forced through the tuning table, against the FP64 CPU oracle.

Tolerances (stated here, used below):
  * plain: ||L L^T - A||_F / ||A||_F (FP64) <= 8 sqrt(n) eps32, and the
    max-norm forward error against the FP64 factor <= 20 x that of
    torch.linalg.cholesky FP32 on the host + 8 eps32;
  * mixed (BF16x3 panel and update products): residual <= 1e-5, forward
    error <= 400 x torch's + 8 eps32;
  * both: the reference's gate (reference/reference.py:137-188), exact zeros
    above the diagonal, input unchanged, every element of L written.
"""
import math
import sys

import numpy as np
import pytest
import torch

from gpumode_cholesky_amd import _lib, autotune, cholesky, harness

pytestmark = pytest.mark.gpu

EPS32 = float(np.finfo(np.float32).eps)
DEV = torch.device("cuda", 0)
FUSED = -2


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _force_fused(n_lo, n_hi):
    autotune.set_table([[n_lo, n_hi, 1, 1 << 30, FUSED]])


def _check(oracle_mod, a_cpu, mode, label):
    a = a_cpu.to(DEV)
    a_copy = a.clone()
    L = cholesky(a, mode=mode)
    torch.cuda.synchronize()
    assert torch.equal(a, a_copy), f"{label}: input was modified"
    ok, msg = harness.check_implementation(a, L)
    assert ok, f"{label}: {msg}"
    Lc = L.cpu().numpy()
    n = a.shape[-1]
    assert (np.triu(Lc, 1) == 0).all(), f"{label}: non-zero above the diagonal"
    L64, info = oracle_mod.potrf64(a_cpu.numpy())
    assert (info == 0).all()
    fro, _, _ = oracle_mod.residual(a_cpu.numpy(), Lc)
    Lt = torch.linalg.cholesky(a_cpu).numpy().astype(np.float64)
    scale = np.abs(L64).reshape(L64.shape[0], -1).max(axis=1)
    e_ours = (np.abs(Lc - L64).reshape(L64.shape[0], -1).max(axis=1) / scale).max()
    e_torch = (np.abs(Lt - L64).reshape(L64.shape[0], -1).max(axis=1) / scale).max()
    _say(f"[fused] {label}: resid={fro.max():.3e} fwd={e_ours:.3e} torch_fwd={e_torch:.3e}")
    if mode == "plain":
        assert fro.max() <= 8.0 * math.sqrt(n) * EPS32, f"{label}: residual {fro.max():.3e}"
        assert e_ours <= 20.0 * e_torch + 8 * EPS32, f"{label}: forward error {e_ours:.3e}"
    else:
        assert fro.max() <= 1e-5, f"{label}: residual {fro.max():.3e}"
        assert e_ours <= 400.0 * e_torch + 8 * EPS32, f"{label}: forward error {e_ours:.3e}"
    return L


@pytest.mark.parametrize("mode", ["plain", "mixed"])
@pytest.mark.parametrize("n", [65, 100, 127, 128, 129, 200, 255, 256, 300, 320, 383, 511, 512, 700, 1100])
def test_fused_against_oracle(oracle_mod, n, mode):
    # Orders around every 128-column step boundary (partial last blocks of
    # 1..127 columns), odd orders (scalar global access), through n = 1100.
    try:
        _force_fused(n, n)
        for case, cond in (("dense", 2), ("spectrum", 5)):
            a = harness.generate_input(3, n, cond, 9100 + n, case, device="cpu")
            _check(oracle_mod, a, mode, f"{mode} n={n} {case}")
    finally:
        autotune.load_table()


@pytest.mark.parametrize("mode", ["plain", "mixed"])
@pytest.mark.parametrize("n", [100, 300, 512])
def test_fused_writes_every_element(n, mode):
    # An output full of NaN comes back identical to a fresh output: the
    # kernel writes the whole of L, zeros above the diagonal included.
    a = harness.generate_input(4, n, 3, 9300 + n, "dense", device="cpu").to(DEV)
    try:
        _force_fused(n, n)
        fresh = cholesky(a, mode=mode)
        out = torch.full_like(a, float("nan"))
        L = cholesky(a, mode=mode, out=out)
        again = cholesky(a, mode=mode)
        torch.cuda.synchronize()
        assert not torch.isnan(L).any().item(), "element not written"
        assert torch.equal(L, fresh), "differs from a fresh output"
        assert torch.equal(again, fresh), "not bitwise repeatable"
    finally:
        autotune.load_table()


@pytest.mark.parametrize("mode", ["plain", "mixed"])
def test_fused_matches_blocked_closely(mode):
    # The same order through the fused and the blocked routes: both factor
    # the same matrices; their results agree to rounding.
    n = 512
    a = harness.generate_input(8, n, 3, 9400, "dense", device="cpu").to(DEV)
    try:
        autotune.set_table([[n, n, 1, 1 << 30, 256]])
        Lb = cholesky(a, mode=mode)
        _force_fused(n, n)
        Lf = cholesky(a, mode=mode)
        torch.cuda.synchronize()
    finally:
        autotune.load_table()
    d = ((Lf - Lb).abs().amax() / Lb.abs().amax()).item()
    _say(f"[fused] {mode} 8x512: max rel difference to the blocked route {d:.3e}")
    assert d <= (1e-5 if mode == "plain" else 1e-4)


@pytest.mark.parametrize("mode", ["plain", "mixed"])
def test_fused_info_on_non_spd(mode):
    # A matrix made indefinite at a chosen minor: info = that minor (1-based);
    # NaN in the failing pivot's column and, through the block's inverse, in
    # the panel columns from it on and the whole trailing factor past its
    # 128-column block; the leading minor before it and the other matrices
    # of the batch finite.
    n = 300
    a = harness.generate_input(3, n, 2, 9500, "dense", device="cpu")
    bad = 170
    a[1, bad - 1, bad - 1] = -1.0
    try:
        _force_fused(n, n)
        L, info = cholesky(a.to(DEV), mode=mode, return_info=True)
        torch.cuda.synchronize()
    finally:
        autotune.load_table()
    info = info.cpu().numpy()
    assert info[0] == 0 and info[2] == 0, info
    assert info[1] == bad, info
    Lc = L.cpu()
    assert torch.isnan(Lc[1, bad - 1:, bad - 1]).all(), "failing pivot's column finite"
    assert torch.isnan(Lc[1, 256:, bad - 1:256]).all(), "panel columns past the failing pivot finite"
    assert torch.isnan(torch.tril(Lc[1, 256:, 256:]) + torch.triu(torch.full((n - 256, n - 256), float("nan")), 1)).all(), \
        "trailing factor finite"
    assert (torch.triu(Lc[1], 1) == 0).all(), "upper triangle not zero"
    assert torch.isfinite(Lc[1, :bad - 1, :bad - 1]).all(), "NaN before the failing pivot"
    assert torch.isfinite(Lc[0]).all() and torch.isfinite(Lc[2]).all()
    assert not torch.isnan(Lc[0]).any() and not torch.isnan(Lc[2]).any()


def test_fused_padded_leading_dimension(oracle_mod):
    # The C ABI with ld > n (row padding filled with NaN in A and L): the
    # padding of A is never read, that of L never written.
    import ctypes
    lib = _lib.load()
    n, ld, batch = 300, 307, 3
    a = harness.generate_input(batch, n, 2, 9600, "dense", device="cpu")
    Ap = torch.full((batch, n, ld), float("nan"))
    Ap[:, :, :n] = a
    Ad = Ap.to(DEV)
    Ld = torch.full((batch, n, ld), float("nan"), device=DEV)
    ws_n = lib.chol_workspace_bytes(batch, n, 0)
    ws = torch.empty((ws_n,), dtype=torch.uint8, device=DEV)
    try:
        _force_fused(n, n)
        assert lib.chol_route(batch, n, 0, None) == 4
        rc = lib.chol_potrf_f32(Ad.data_ptr(), Ld.data_ptr(), batch, n, n * ld, ld, 0, None, ws.data_ptr(), ws_n,
                                ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
        torch.cuda.synchronize()
    finally:
        autotune.load_table()
    assert rc == 0, _lib.last_error()
    Lc = Ld.cpu()
    assert torch.isnan(Lc[:, :, n:]).all(), "row padding of L written"
    L = Lc[:, :, :n].contiguous().numpy()
    assert (np.triu(L, 1) == 0).all()
    fro, _, _ = oracle_mod.residual(a.numpy(), L)
    assert fro.max() <= 8.0 * math.sqrt(n) * EPS32


@pytest.mark.parametrize("mode", ["plain", "mixed"])
def test_fused_full_shape_640x512(mode):
    # The benchmark shape through the fused route at full size: the gate,
    # the FP64 residual and the oracle's forward error on sampled matrices.
    from gpumode_cholesky_amd import harness as h
    a = h.config_input(640, 512, 5120, device=DEV)
    try:
        _force_fused(512, 512)
        L = cholesky(a, mode=mode)
        torch.cuda.synchronize()
    finally:
        autotune.load_table()
    ok, msg = h.check_implementation(a, L)
    assert ok, msg
    fro = h.fp64_residual(a, L, chunk=64)
    _say(f"[fused] {mode} 640x512: resid={fro:.3e} {msg}")
    assert fro <= (8.0 * math.sqrt(512) * EPS32 if mode == "plain" else 1e-5)
