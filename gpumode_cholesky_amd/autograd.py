"""Differentiable Cholesky on the HIP kernels.

``cholesky_ad(A)`` returns ``L`` like ``cholesky`` but records a
``torch.autograd.Function`` whose backward is the hand-written adjoint kernel
(``chol_backward_f32``):

    Abar = 1/2 * sym(L^-T Phi(L^T Lbar) L^-1),  Phi = tril with halved diagonal,

the symmetric gradient torch's autograd gives for ``torch.linalg.cholesky``.
With ``store_inverse=True`` the forward also forms ``L^-1``
(``chol_tri_inverse_f32``) and keeps it for the backward; otherwise the
backward recomputes it (less memory held between the passes).

The reference (submission.py) has no backward; the formula is pinned by
FP64 finite differences and torch autograd in tests/test_backward_gpu.py.
"""
from __future__ import annotations

import torch

from . import _lib
from .cholesky import _raw_stream, _validate, cholesky


def _batched(x: torch.Tensor):
    n = x.shape[-1]
    return x.reshape(-1, n, n).contiguous(), n


def tri_inverse(L: torch.Tensor) -> torch.Tensor:
    """``L^-1`` of lower-triangular FP32 matrices ([..., n, n])."""
    _validate(L)
    L3, n = _batched(L)
    b = L3.shape[0]
    lib = _lib.load()
    X = torch.empty_like(L3)
    nws = lib.chol_tri_inverse_workspace_bytes(b, n)
    ws = torch.empty((max(nws, 1),), dtype=torch.uint8, device=L.device)
    with torch.cuda.device(L.device):
        rc = lib.chol_tri_inverse_f32(L3.data_ptr(), X.data_ptr(), b, n, ws.data_ptr(), nws,
                                      _raw_stream(L.device.index))
    _lib.check(rc, "chol_tri_inverse_f32")
    return X.view(L.shape)


def cholesky_backward(L: torch.Tensor, Lbar: torch.Tensor, Linv: torch.Tensor | None = None) -> torch.Tensor:
    """Adjoint of the factorisation: ``Abar`` (symmetric) from ``L`` and the
    upstream gradient ``Lbar`` (its strictly upper part does not matter)."""
    _validate(L)
    L3, n = _batched(L)
    G3, _ = _batched(Lbar.to(torch.float32))
    b = L3.shape[0]
    lib = _lib.load()
    Abar = torch.empty_like(L3)
    have = Linv is not None
    X3 = _batched(Linv)[0] if have else None
    nws = lib.chol_backward_workspace_bytes(b, n)
    ws = torch.empty((max(nws, 1),), dtype=torch.uint8, device=L.device)
    with torch.cuda.device(L.device):
        rc = lib.chol_backward_f32(L3.data_ptr(), G3.data_ptr(), Abar.data_ptr(), b, n,
                                   X3.data_ptr() if have else None, 1 if have else 0,
                                   ws.data_ptr(), nws, _raw_stream(L.device.index))
    _lib.check(rc, "chol_backward_f32")
    return Abar.view(L.shape)


class CholeskyFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, A, mode="plain", store_inverse=False):
        L = cholesky(A, mode=mode)
        if store_inverse:
            ctx.save_for_backward(L, tri_inverse(L))
        else:
            ctx.save_for_backward(L)
        ctx.store_inverse = store_inverse
        return L

    @staticmethod
    def backward(ctx, gL):
        if ctx.store_inverse:
            L, Linv = ctx.saved_tensors
        else:
            (L,) = ctx.saved_tensors
            Linv = None
        return cholesky_backward(L, gL, Linv), None, None


def cholesky_ad(A: torch.Tensor, mode: str = "plain", store_inverse: bool = False) -> torch.Tensor:
    """``cholesky(A)`` with the hand-written backward attached."""
    return CholeskyFunction.apply(A, mode, store_inverse)
