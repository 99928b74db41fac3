"""GPU parity of the triangular inverse and the hand-written backward.

Pins (no reference counterpart exists; the reference has no backward):
  * tri_inverse: against the FP64 oracle inverse, relative max error
    <= INV_TOL * cond-scale, and L @ Linv = I;
  * backward: against torch autograd of torch.linalg.cholesky in FP64 on the
    host, and against central finite differences in FP64 (h = 1e-6) along
    random symmetric directions; relative error <= BWD_TOL (1e-4, the
    target the benchmark states for the backward).
"""
import math

import numpy as np
import pytest
import torch

from gpumode_cholesky_amd import harness
from gpumode_cholesky_amd.autograd import cholesky_ad, cholesky_backward, tri_inverse

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda", 0)
BWD_TOL = 1e-4


def _spd(batch, n, cond, seed):
    return harness.spectral_input(batch, n, cond, seed, device="cpu")


@pytest.mark.parametrize("n", [1, 5, 64, 65, 100, 128, 200, 300, 1024])
def test_tri_inverse_against_oracle(oracle_mod, n):
    a = _spd(3, n, 1e2, 10 + n)
    L64, _ = oracle_mod.potrf64(a.numpy())
    L = torch.from_numpy(L64.astype(np.float32))
    X = tri_inverse(L.to(DEV)).cpu().numpy()
    X64 = oracle_mod.tri_inverse64(L.numpy().astype(np.float64))
    err = np.abs(X - X64).max() / np.abs(X64).max()
    print(f"[tri_inverse] n={n}: rel err {err:.3e}", flush=True)
    assert err < 1e-5
    assert (np.triu(X, 1) == 0).all()


def _torch_grad64(a: torch.Tensor, gL: torch.Tensor) -> torch.Tensor:
    A = a.double().clone().requires_grad_(True)
    L = torch.linalg.cholesky(A)
    (L * gL.double()).sum().backward()
    return A.grad


@pytest.mark.parametrize("batch,n", [(4, 4), (4, 32), (3, 64), (2, 100), (2, 128), (2, 200), (1, 300)])
@pytest.mark.parametrize("store", [False, True])
def test_backward_against_torch_autograd(batch, n, store):
    a = _spd(batch, n, 1e2, 100 + n)
    gL = torch.tril(torch.randn(batch, n, n, generator=torch.Generator().manual_seed(n)))
    A = a.to(DEV).requires_grad_(True)
    L = cholesky_ad(A, store_inverse=store)
    (L * gL.to(DEV)).sum().backward()
    g = A.grad.cpu().double()
    ref = _torch_grad64(a, gL)
    err = (g - ref).abs().max().item() / ref.abs().max().item()
    print(f"[backward] {batch}x{n} store={store}: rel err vs torch fp64 {err:.3e}", flush=True)
    assert err <= BWD_TOL
    assert torch.allclose(g, g.transpose(-1, -2))


def test_backward_store_and_recompute_agree():
    a = _spd(4, 160, 1e2, 7).to(DEV)
    L = torch.linalg.cholesky(a.cpu()).to(DEV)
    gL = torch.tril(torch.randn(4, 160, 160, generator=torch.Generator().manual_seed(1))).to(DEV)
    g1 = cholesky_backward(L, gL)
    g2 = cholesky_backward(L, gL, tri_inverse(L))
    assert torch.equal(g1, g2)


def test_backward_finite_differences(oracle_mod):
    # Config: 256 x 128, spectral cond 1e2; 8 sampled matrices, central FD in
    # FP64 along random symmetric directions E: <Abar, E> vs
    # (f(A + hE) - f(A - hE)) / 2h with f(A) = sum(Lbar * chol64(A)).
    a = _spd(256, 128, 1e2, 2024)
    gL = torch.tril(torch.randn(256, 128, 128, generator=torch.Generator().manual_seed(5)))
    A = a.to(DEV).requires_grad_(True)
    L = cholesky_ad(A)
    (L * gL.to(DEV)).sum().backward()
    g = A.grad.cpu().double().numpy()
    rng = np.random.default_rng(0)
    h = 1e-6
    worst = 0.0
    for b in np.linspace(0, 255, 8).astype(int):
        a64 = a[b].double().numpy()
        w = gL[b].double().numpy()
        for _ in range(2):
            E = rng.standard_normal((128, 128))
            E = 0.5 * (E + E.T)
            fp = (w * np.linalg.cholesky(a64 + h * E)).sum()
            fm = (w * np.linalg.cholesky(a64 - h * E)).sum()
            fd = (fp - fm) / (2 * h)
            an = (g[b] * E).sum()
            rel = abs(an - fd) / max(abs(fd), 1e-12)
            worst = max(worst, rel)
    print(f"[backward-fd] worst rel err {worst:.3e}", flush=True)
    assert worst <= BWD_TOL


def test_backward_finite_differences_8x2048():
    # Config: 8 x 2048, spectral cond 1e2; all 8 matrices, central FD in FP64
    # (h = 1e-6) on the host along two random symmetric directions each.
    # Relative error of <Abar, E> against the FD quotient <= BWD_TOL.
    a = harness.spectral_input(8, 2048, 1e2, 2048, device=DEV)
    gL = torch.tril(torch.randn(8, 2048, 2048, device=DEV, generator=torch.Generator(device=DEV).manual_seed(9)))
    A = a.clone().requires_grad_(True)
    L = cholesky_ad(A)
    (L * gL).sum().backward()
    g = A.grad.cpu().double().numpy()
    a_h = a.cpu().double().numpy()
    w_h = gL.cpu().double().numpy()
    del A, L, gL
    rng = np.random.default_rng(1)
    h = 1e-6
    worst = 0.0
    for b in range(8):
        for _ in range(2):
            E = rng.standard_normal((2048, 2048))
            E = 0.5 * (E + E.T)
            fp = (w_h[b] * np.linalg.cholesky(a_h[b] + h * E)).sum()
            fm = (w_h[b] * np.linalg.cholesky(a_h[b] - h * E)).sum()
            fd = (fp - fm) / (2 * h)
            an = (g[b] * E).sum()
            rel = abs(an - fd) / max(abs(fd), 1e-12)
            print(f"[backward-fd-2048] matrix {b}: <Abar,E>={an:.6e} fd={fd:.6e} rel={rel:.3e}", flush=True)
            worst = max(worst, rel)
    assert worst <= BWD_TOL


def test_backward_large_config():
    # 8 x 2048, cond 1e2: two sampled matrices against torch fp64 autograd.
    a = harness.spectral_input(8, 2048, 1e2, 77, device=DEV)
    gL = torch.tril(torch.randn(8, 2048, 2048, device=DEV, generator=torch.Generator(device=DEV).manual_seed(3)))
    A = a.clone().requires_grad_(True)
    L = cholesky_ad(A)
    (L * gL).sum().backward()
    for b in (0, 7):
        ref = _torch_grad64(a[b].cpu(), gL[b].cpu())
        err = (A.grad[b].cpu().double() - ref).abs().max().item() / ref.abs().max().item()
        print(f"[backward-2048] matrix {b}: rel err {err:.3e}", flush=True)
        assert err <= BWD_TOL


def test_inverse_and_backward_beyond_grid_limit():
    # 32769 matrices of 256: the batched products of the inverse put
    # matrices x diagonal pairs (and the backward's passes the matrices) on
    # a grid dimension limited to 65535, so the launches split; checked on
    # sampled matrices (both sides of each split) against FP64 on the host.
    batch, n = 32769, 256
    a = harness.spectral_input(8, n, 1e2, 4321, device=DEV)
    idx = [0, 8191, 16383, 16384, 32767, 32768]
    A = a[torch.arange(batch, device=DEV) % 8].contiguous()
    del a
    L = cholesky_ad(A)  # forward: blocked route for 32769 x 256
    gL = torch.tril(torch.randn(batch, n, n, device=DEV, generator=torch.Generator(device=DEV).manual_seed(2)))
    X = tri_inverse(L)
    G = cholesky_backward(L, gL)
    torch.cuda.synchronize()
    for b in idx:
        Lb = L[b].cpu().double()
        Xr = torch.linalg.inv(Lb)
        e_inv = (X[b].cpu().double() - Xr).abs().max().item() / Xr.abs().max().item()
        ref = _torch_grad64(A[b].cpu(), gL[b].cpu())
        e_bwd = (G[b].cpu().double() - ref).abs().max().item() / ref.abs().max().item()
        print(f"[beyond-65535] matrix {b}: inverse {e_inv:.3e} backward {e_bwd:.3e}", flush=True)
        assert e_inv < 1e-5 and e_bwd <= BWD_TOL
