"""GPU parity of the HIP path (through the C ABI) against the FP64 CPU oracle
and against torch.linalg.cholesky on the same device.

Tolerances (FP32 arithmetic, stated here and used below):
  * reconstruction: ||L L^T - A||_F / ||A||_F (FP64) <= RESID_C * sqrt(n) * eps32
  * forward error:  ||L - L64||_max / ||L64||_max <= FWD_C * e_torch + FWD_ABS,
    where e_torch is the same error of torch.linalg.cholesky in FP32 (the
    reference's own ref_kernel, reference/reference.py:129-130, run on the
    host) and L64 the FP64 oracle factor;
  * the reference's gate (reference/reference.py:137-188) must pass;
  * exact zeros above the diagonal, positive diagonal, input unchanged.
"""
import math
import sys

import numpy as np
import pytest
import torch

from gpumode_cholesky_amd import _lib, cholesky, custom_kernel
from gpumode_cholesky_amd import harness

pytestmark = pytest.mark.gpu

EPS32 = float(np.finfo(np.float32).eps)
RESID_C = 8.0
FWD_C = 20.0
FWD_ABS = 8 * EPS32
DEV = torch.device("cuda", 0)


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _check_against_oracle(oracle_mod, a_cpu, label):
    a = a_cpu.to(DEV)
    a_copy = a.clone()
    L = custom_kernel(a)
    torch.cuda.synchronize()
    assert torch.equal(a, a_copy), "input was modified"
    ok, msg = harness.check_implementation(a, L)
    assert ok, f"{label}: {msg}"
    Lc = L.cpu().numpy()
    n = a.shape[-1]
    assert (np.triu(Lc, 1) == 0).all(), "non-zero above the diagonal"
    L64, info = oracle_mod.potrf64(a_cpu.numpy())
    assert (info == 0).all()
    fro, _, _ = oracle_mod.residual(a_cpu.numpy(), Lc)
    # torch.linalg.cholesky in FP32 on the host (LAPACK spotrf): the
    # reference's own path on a CPU box (submission.py:4268).
    Lt = torch.linalg.cholesky(a_cpu).numpy().astype(np.float64)
    scale = np.abs(L64).reshape(L64.shape[0], -1).max(axis=1)
    e_ours = (np.abs(Lc - L64).reshape(L64.shape[0], -1).max(axis=1) / scale).max()
    e_torch = (np.abs(Lt - L64).reshape(L64.shape[0], -1).max(axis=1) / scale).max()
    _say(f"[parity] {label}: resid={fro.max():.3e} fwd={e_ours:.3e} torch_fwd={e_torch:.3e} {msg}")
    assert fro.max() <= RESID_C * math.sqrt(n) * EPS32, f"{label}: residual {fro.max():.3e}"
    assert e_ours <= FWD_C * e_torch + FWD_ABS, f"{label}: forward error {e_ours:.3e} vs torch {e_torch:.3e}"
    return L


SMALL_N = [1, 2, 3, 5, 8, 16, 17, 31, 32, 33, 48, 63, 64]


@pytest.mark.parametrize("case,cond", harness.FAMILIES)
@pytest.mark.parametrize("n", SMALL_N)
def test_small_families(oracle_mod, n, case, cond):
    a = harness.generate_input(7, n, cond, 1000 + n, case, device="cpu")
    _check_against_oracle(oracle_mod, a, f"small n={n} {case}")


BLOCKED_N = [65, 96, 127, 128, 129, 200, 256, 300, 320, 512, 700]


@pytest.mark.parametrize("case,cond", [("dense", 2), ("spectrum", 5), ("rowscale", 4), ("tridiagonal", 1)])
@pytest.mark.parametrize("n", BLOCKED_N)
def test_blocked_families(oracle_mod, n, case, cond):
    a = harness.generate_input(3, n, cond, 2000 + n, case, device="cpu")
    _check_against_oracle(oracle_mod, a, f"blocked n={n} {case}")


@pytest.mark.parametrize("nbo", [128, -1])
@pytest.mark.parametrize("n", [65, 100, 127, 128, 129, 200, 256])
def test_small_order_routes_forced(oracle_mod, n, nbo):
    # Both routes for 64 < n <= 256, forced through the table.  Blocked (the
    # default below a batch of 8192): one 128-block (n <= 128) is staged from
    # A by the diagonal-block kernel with no copy before it; n > 128 ends
    # with a block that forms no inverse.  -1: the whole-matrix LDS kernel
    # (the default from a batch of 8192).
    from gpumode_cholesky_amd import autotune
    try:
        autotune.set_table([[n, n, 1, 1 << 20, nbo]])
        for case, cond in (("dense", 2), ("spectrum", 5)):
            a = harness.generate_input(3, n, cond, 5000 + n, case, device="cpu")
            _check_against_oracle(oracle_mod, a, f"route nbo={nbo} n={n} {case}")
    finally:
        autotune.load_table()


@pytest.mark.parametrize("mode", ["plain", "mixed"])
@pytest.mark.parametrize("n", [32, 64, 100, 128, 200, 256, 383, 512, 1001])
def test_output_written_whole(n, mode):
    # The kernels write every element of L, zeros above the diagonal
    # included (there is no separate zero fill), so an output buffer full of
    # NaN comes back as the same factor as a fresh one.  n <= 256 also runs
    # with the blocked route and (64 < n) the LDS kernel forced.
    from gpumode_cholesky_amd import autotune
    a = harness.generate_input(3, n, 3, 7100 + n, "dense", device="cpu").to(DEV)
    ref = cholesky(a, mode=mode)
    routes = [None] if n > 256 else [None, [[n, n, 1, 1 << 20, 128]]]
    if 64 < n <= 256:
        routes.append([[n, n, 1, 1 << 20, -1]])
    try:
        for table in routes:
            if table is not None:
                autotune.set_table(table)
            fresh = cholesky(a, mode=mode)
            out = torch.full_like(a, float("nan"))
            L = cholesky(a, mode=mode, out=out)
            torch.cuda.synchronize()
            assert L.data_ptr() == out.data_ptr()
            assert torch.count_nonzero(torch.triu(L, 1)).item() == 0, f"n={n} route={table}: upper not zero"
            assert not torch.isnan(L).any().item(), f"n={n} route={table}: element not written"
            assert torch.equal(L, fresh), f"n={n} route={table}: differs from a fresh output"
            if table is None:
                assert torch.equal(L, ref)
    finally:
        autotune.load_table()


@pytest.mark.parametrize("batch,n", [(4096, 32), (1024, 64), (256, 128), (64, 256), (16, 512), (4, 1024)])
def test_config_shapes_against_oracle(oracle_mod, batch, n):
    # The benchmark workload generator at its real shape; the oracle checks a
    # sampled subset of the batch (the whole batch for the smaller n).
    a = harness.config_input(batch, n, 77 + n, device="cpu")
    L = custom_kernel(a.to(DEV)).cpu()
    idx = np.linspace(0, batch - 1, min(batch, 16)).astype(int)
    sub = a[idx]
    L64, _ = oracle_mod.potrf64(sub.numpy())
    fro, _, _ = oracle_mod.residual(sub.numpy(), L[idx].numpy())
    e = np.abs(L[idx].numpy() - L64).max() / np.abs(L64).max()
    _say(f"[config] {batch}x{n}: resid={fro.max():.3e} fwd={e:.3e}")
    assert fro.max() <= RESID_C * math.sqrt(n) * EPS32
    ok, msg = harness.check_implementation(a.to(DEV), L.to(DEV))
    assert ok, msg


MIXED_RESID = 1.0e-5  # ||LL^T - A||_F / ||A||_F target of the mixed form
FWD_MIXED_C = 400.0  # mixed form: BF16x3 products carry ~2^-16 where FP32 carries ~2^-24


@pytest.mark.parametrize("mode", ["plain", "mixed"])
@pytest.mark.parametrize("spec", harness.BENCH_SPECS, ids=lambda s: f"{s['batch']}x{s['n']}")
def test_bench_shapes_gate(spec, mode):
    # Full benchmark shapes, inputs generated on the GPU: the reference gate,
    # the FP64 residual (<= 8 sqrt(n) eps32 plain, <= 1e-5 mixed), exact zeros
    # above the diagonal, and for n <= 16384 the max-norm forward error of
    # matrix 0 against its FP64 factor (LAPACK dpotrf on the host), bounded
    # by that of FP32 LAPACK (torch.linalg.cholesky on the host, the
    # reference's own CPU path) on the same matrix: <= 20 x (plain) or
    # <= 400 x (mixed) of it, + 8 eps32.
    b, n = spec["batch"], spec["n"]
    _say(f"[bench-shape] start {b}x{n} {mode}")
    a = harness.config_input(b, n, spec["seed"], device=DEV)
    L = cholesky(a, mode=mode)
    torch.cuda.synchronize()
    ok, msg = harness.check_implementation(a, L)
    assert ok, f"{b}x{n}: {msg}"
    resid = harness.fp64_residual(a, L)
    line = f"[bench-shape] {b}x{n} {mode}: {msg} fp64_resid={resid:.3e}"
    assert resid <= (MIXED_RESID if mode == "mixed" else RESID_C * math.sqrt(n) * EPS32), line
    assert torch.count_nonzero(torch.triu(L, 1)).item() == 0
    if n <= 16384:
        a0 = a[0].cpu()
        L64 = np.linalg.cholesky(a0.double().numpy())
        Lt = torch.linalg.cholesky(a0).double().numpy()
        scale = np.abs(L64).max()
        e_ours = np.abs(L[0].cpu().double().numpy() - L64).max() / scale
        e_torch = np.abs(Lt - L64).max() / scale
        line += f" fwd={e_ours:.3e} torch_fwd={e_torch:.3e}"
        c = FWD_MIXED_C if mode == "mixed" else FWD_C
        assert e_ours <= c * e_torch + FWD_ABS, line
    _say(line)
    del a, L
    torch.cuda.empty_cache()




@pytest.mark.parametrize("batch,n,case", [(3, 300, "dense"), (2, 512, "spectrum"), (2, 1024, "dense"),
                                          (2, 1024, "spectrum"), (1, 2048, "rowscale")])
def test_mixed_against_oracle(oracle_mod, batch, n, case):
    # BF16x3 trailing updates: residual within MIXED_RESID of the FP64
    # oracle's factorisation target, the reference gate passes, and the
    # factor stays within 1e-3 (relative, max-norm) of the FP64 oracle factor.
    cond = {"dense": 2, "spectrum": 5, "rowscale": 4}[case]
    a = harness.generate_input(batch, n, cond, 3000 + n, case, device="cpu")
    L = cholesky(a.to(DEV), mode="mixed")
    ok, msg = harness.check_implementation(a.to(DEV), L)
    assert ok, msg
    Lc = L.cpu().numpy()
    fro, _, _ = oracle_mod.residual(a.numpy(), Lc)
    L64, _ = oracle_mod.potrf64(a.numpy())
    e = np.abs(Lc - L64).max() / np.abs(L64).max()
    _say(f"[mixed] {batch}x{n} {case}: resid={fro.max():.3e} fwd={e:.3e} {msg}")
    assert fro.max() <= MIXED_RESID
    assert e <= 1e-3
    assert (np.triu(Lc, 1) == 0).all()


@pytest.mark.parametrize("nbo", [384, 640])
@pytest.mark.parametrize("mode", ["plain", "mixed"])
def test_outer_block_widths_not_powers_of_two(oracle_mod, mode, nbo):
    # Outer blocks of three or five 128-column steps through the tuning
    # table, n = 1100 leaving a partial last block (and a last diagonal block
    # of 76 columns): look-ahead bulk and narrow updates at these widths.
    from gpumode_cholesky_amd import autotune
    n = 1100
    a = harness.generate_input(2, n, 3, 7300 + nbo, "dense", device="cpu")
    try:
        autotune.set_table([[n, n, 1, 1 << 20, nbo]])
        L = cholesky(a.to(DEV), mode=mode)
        torch.cuda.synchronize()
    finally:
        autotune.load_table()
    ok, msg = harness.check_implementation(a.to(DEV), L)
    assert ok, msg
    Lc = L.cpu().numpy()
    fro, _, _ = oracle_mod.residual(a.numpy(), Lc)
    _say(f"[nbo] {mode} n={n} nbo={nbo}: resid={fro.max():.3e}")
    assert fro.max() <= (MIXED_RESID if mode == "mixed" else RESID_C * math.sqrt(n) * EPS32)
    assert (np.triu(Lc, 1) == 0).all()


@pytest.mark.parametrize("mode", ["plain", "mixed"])
def test_unaligned_rows_wide_updates(oracle_mod, mode):
    # n = 1001 (rows not 16-byte aligned) and 96 matrices: the outer-block
    # updates have >= 256 tiles of 256 x 256 and K = 256, the 8-wave 256-tile
    # SYRK with scalar row loads.  Checked on 8 sampled matrices.
    batch, n = 96, 1001
    a = harness.generate_input(batch, n, 2, 4242, device=DEV)
    L = cholesky(a, mode=mode)
    ok, msg = harness.check_implementation(a, L)
    assert ok, msg
    idx = np.linspace(0, batch - 1, 8).astype(int)
    sub = a[idx].cpu().numpy()
    Lc = L[idx].cpu().numpy()
    fro, _, _ = oracle_mod.residual(sub, Lc)
    _say(f"[unaligned-wide] {batch}x{n} {mode}: resid={fro.max():.3e} {msg}")
    assert fro.max() <= (MIXED_RESID if mode == "mixed" else RESID_C * math.sqrt(n) * EPS32)
    assert (np.triu(Lc, 1) == 0).all()


@pytest.mark.parametrize("mode", ["plain", "mixed"])
@pytest.mark.parametrize("n,ld", [(300, 304), (640, 644), (640, 641)])
def test_padded_leading_dimension(oracle_mod, mode, n, ld):
    # Through the C ABI with rows padded to ld: a multiple of 4 keeps the
    # float4 paths (and, mixed, the next block's update inside the
    # diagonal-block kernel), an odd ld takes the scalar paths.  The padding
    # of A is NaN (never read) and that of L a sentinel (never written).
    batch = 3
    a = harness.generate_input(batch, n, 2, 5100 + n, device="cpu")
    A = torch.full((batch, n, ld), float("nan"))
    A[:, :, :n] = a
    L = torch.full((batch, n, ld), -7.0)
    A, L = A.to(DEV), L.to(DEV)
    lib = _lib.load()
    m = 1 if mode == "mixed" else 0
    ws_n = lib.chol_workspace_bytes(batch, n, m)
    ws = torch.empty((max(ws_n, 1),), dtype=torch.uint8, device=DEV)
    _say(f"[padded-ld] start {batch}x{n} ld={ld} {mode}")
    rc = lib.chol_potrf_f32(A.data_ptr(), L.data_ptr(), batch, n, n * ld, ld, m, None, ws.data_ptr(), ws_n,
                            torch.cuda.current_stream(DEV).cuda_stream)
    torch.cuda.synchronize()
    assert rc == 0, lib.chol_last_error()
    Lc = L.cpu()
    assert (Lc[:, :, n:] == -7.0).all()
    Lf = Lc[:, :, :n].contiguous().numpy()
    fro, _, _ = oracle_mod.residual(a.numpy(), Lf)
    _say(f"[padded-ld] {batch}x{n} ld={ld} {mode}: resid={fro.max():.3e}")
    assert fro.max() <= (MIXED_RESID if mode == "mixed" else RESID_C * math.sqrt(n) * EPS32)
    assert (np.triu(Lf, 1) == 0).all()


@pytest.mark.parametrize("n", [7, 32, 64, 100, 300])
def test_non_spd_sets_info_and_nan(n):
    a = harness.generate_input(3, n, 2, 5, device="cpu")
    k = n // 2
    a[1, k, k] = -1.0  # the leading minor of order k+1 is indefinite
    L, info = cholesky(a.to(DEV), return_info=True)
    info = info.cpu()
    assert info[0].item() == 0 and info[2].item() == 0
    assert info[1].item() == k + 1
    Lc = L.cpu()
    assert torch.isfinite(Lc[0]).all() and torch.isfinite(Lc[2]).all()
    assert torch.isnan(Lc[1, k, k])


@pytest.mark.parametrize("n", [24, 64, 160])
def test_strided_and_transposed_inputs(oracle_mod, n):
    base = harness.generate_input(4, n, 2, 9, device="cpu").to(DEV)
    ref = custom_kernel(base.contiguous())
    # non-contiguous batch (every other matrix of a larger buffer)
    big = torch.zeros((8, n, n), device=DEV)
    big[::2] = base
    out = custom_kernel(big[::2])
    assert torch.equal(out, ref)
    # transposed view of a symmetric matrix: same lower triangle
    out_t = custom_kernel(base.transpose(-1, -2))
    assert torch.equal(out_t, ref)
    # 2-D input
    out2 = custom_kernel(base[1])
    assert out2.shape == (n, n) and torch.equal(out2, ref[1])


@pytest.mark.parametrize("n", [32, 64, 128, 512])
def test_upper_triangle_of_input_is_ignored(n):
    a = harness.generate_input(2, n, 2, 3, device="cpu").to(DEV)
    garbage = a.clone()
    iu = torch.triu_indices(n, n, 1, device=DEV)
    garbage[:, iu[0], iu[1]] = float("nan")
    assert torch.equal(custom_kernel(garbage), custom_kernel(a))


@pytest.mark.parametrize("batch,n", [(4096, 32), (1024, 64), (64, 256), (4, 1024)])
def test_bitwise_determinism(batch, n):
    a = harness.config_input(batch, n, 5, device=DEV)
    outs = [custom_kernel(a) for _ in range(3)]
    assert torch.equal(outs[0], outs[1]) and torch.equal(outs[1], outs[2])


def test_empty_batch_and_n1():
    a = torch.empty((0, 5, 5), device=DEV)
    assert custom_kernel(a).shape == (0, 5, 5)
    a = torch.full((3, 1, 1), 9.0, device=DEV)
    assert torch.equal(custom_kernel(a), torch.full((3, 1, 1), 3.0, device=DEV))


@pytest.mark.parametrize("mode", ["plain", "mixed"])
@pytest.mark.parametrize("batch,n", [(2, 2048), (1, 4096), (1, 8192)])
def test_large_spectral_cond_1e5(oracle_mod, batch, n, mode):
    # The large-n workload's cond = 1e5 spectral variant (Q diag(lambda) Q^T,
    # lambda log-spaced over 5 decades): FP64 residual <= 1e-5 (mixed) or
    # <= 8 sqrt(n) eps32 (plain), the reference gate, and at 2 x 2048 the
    # max-norm forward error of matrix 1 against the FP64 oracle factor,
    # bounded by that of torch.linalg.cholesky in FP32 on the host (LAPACK):
    # <= 20 x (plain) or <= 400 x (mixed) of it, + 8 eps32.
    _say(f"[cond1e5] start {batch}x{n} {mode}")
    a = harness.spectral_input(batch, n, 1e5, 6100 + n, device=DEV)
    L = cholesky(a, mode=mode)
    torch.cuda.synchronize()
    ok, msg = harness.check_implementation(a, L)
    assert ok, msg
    resid = harness.fp64_residual(a, L)
    line = f"[cond1e5] {batch}x{n} {mode}: {msg} fp64_resid={resid:.3e}"
    assert resid <= (MIXED_RESID if mode == "mixed" else RESID_C * math.sqrt(n) * EPS32), line
    if n == 2048:
        a1 = a[1:2].cpu()
        L64, info = oracle_mod.potrf64(a1.numpy())
        assert info[0] == 0
        Lt = torch.linalg.cholesky(a1).numpy().astype(np.float64)
        scale = np.abs(L64).max()
        e_ours = np.abs(L[1:2].cpu().numpy() - L64).max() / scale
        e_torch = np.abs(Lt - L64).max() / scale
        line += f" fwd={e_ours:.3e} torch_fwd={e_torch:.3e}"
        c = FWD_MIXED_C if mode == "mixed" else FWD_C
        assert e_ours <= c * e_torch + FWD_ABS, line
    _say(line)
    del a, L
    torch.cuda.empty_cache()
