"""Host side of the batched Cholesky path: torch tensors in, C ABI calls out.

``custom_kernel(data) -> L`` is the drop-in for the reference's entry point
(submission.py:4101-4268, called by reference/eval.py:107, :161, :234): a
``batch x n x n`` FP32 SPD device tensor in, a NEW lower-triangular FP32
tensor with positive diagonal out, the input left untouched
(reference/eval.py:180 reuses inputs uncloned).

Every call goes to ``libchol_amd.so``; there is no CPU or library fallback.
Non-SPD input follows the reference's ``cholesky_ex(check_errors=False)``
(reference/reference.py:129-130): no exception, NaN in the factor from the
failing pivot on, and ``info`` (when asked for) holding the first failing
minor.
"""
from __future__ import annotations

import math

import torch

from . import _lib

_MODES = {"plain": _lib.MODE_PLAIN, "mixed": _lib.MODE_MIXED, "fp8": _lib.MODE_FP8}

try:  # raw hipStream_t of the current stream, without building a Stream object
    _raw_stream = torch._C._cuda_getCurrentRawStream
except AttributeError:  # pragma: no cover - older torch
    def _raw_stream(index):
        return torch.cuda.current_stream(index).cuda_stream


def _validate(A: torch.Tensor) -> torch.Tensor:
    if not isinstance(A, torch.Tensor):
        raise TypeError("input must be a torch.Tensor")
    if A.dtype != torch.float32:
        raise TypeError(f"input dtype must be torch.float32, got {A.dtype}")
    if A.device.type != "cuda":
        raise ValueError(
            f"input must be on a HIP device (got {A.device}); this engine has "
            "no CPU path")
    if A.dim() < 2 or A.shape[-1] != A.shape[-2]:
        raise ValueError(f"input must be [..., n, n], got {tuple(A.shape)}")
    return A


def workspace_bytes(batch: int, n: int, mode: str = "plain") -> int:
    return int(_lib.load().chol_workspace_bytes(batch, n, _MODES[mode]))


# Matrices the mixed form handed to the plain form (fallback=True), since load.
MIXED_FALLBACKS = 0


def cholesky(A: torch.Tensor, mode: str = "plain", return_info: bool = False,
             out: torch.Tensor | None = None, fallback: bool = False):
    """Lower Cholesky factor of each matrix of ``A`` ([..., n, n] FP32).

    mode: ``"plain"`` (all-FP32) or ``"mixed"`` (BF16x3 split products in the
    panel solves and trailing updates of the blocked path, which runs for
    n > 256 and for 64 < n <= 256 where the tuning table picks it; the
    register and LDS kernels are the same in both forms), or ``"fp8"``: the
    mixed form with the bulk trailing updates of orders >= 16384 in FP8 E4M3
    (rows scaled by powers of two from A's diagonal), the reference's MXFP8
    large-n form; its residual (~1e-3..1e-2) passes the reference's gate
    but not the 1e-5 of the mixed form.
    fallback (mixed and fp8 forms): BF16x3 products carry ~2^-16 relative error, so on
    ill-conditioned input (cond ~1e5 and beyond, e.g. a low-rank matrix plus
    1e-4 I) a pivot can turn non-positive.  With fallback=True the call reads
    info back (one host synchronisation) and refactors those matrices in the
    plain form, as the reference's isfinite fallback chain does
    (submission.py:4158-4164, :4268); without it they come back as NaN with
    info set, like any non-SPD input.
    Returns ``L`` or ``(L, info)`` with ``info`` int32 [...] on the device.
    """
    m = _MODES.get(mode)
    if m is None:
        raise ValueError(f"mode must be one of {sorted(_MODES)}, got {mode!r}")
    _validate(A)
    if fallback and m != _lib.MODE_PLAIN:
        return _mixed_with_fallback(A, return_info, out, mode)
    lib = _lib.load()
    shape = A.shape
    n = shape[-1]
    # The C ABI gives A and L one layout; L is contiguous, so A must be too.
    A3 = A if A.is_contiguous() else A.contiguous()
    batch = math.prod(shape[:-2])
    if out is None:
        L = torch.empty(shape, dtype=torch.float32, device=A.device)
    else:
        if out.shape != shape or out.dtype != torch.float32 or out.device != A.device \
                or not out.is_contiguous():
            raise ValueError("out must be a contiguous FP32 tensor shaped like A")
        L = out
    info = torch.empty(shape[:-2], dtype=torch.int32, device=A.device) if return_info else None
    ws_n = lib.chol_workspace_bytes(batch, n, m)
    ws = torch.empty((ws_n,), dtype=torch.uint8, device=A.device) if ws_n else None
    dev = A.device.index
    if dev != torch.cuda.current_device():
        with torch.cuda.device(dev):
            rc = lib.chol_potrf_f32(A3.data_ptr(), L.data_ptr(), batch, n, n * n, max(n, 1), m,
                                    info.data_ptr() if info is not None else None,
                                    ws.data_ptr() if ws is not None else None, ws_n, _raw_stream(dev))
    else:
        rc = lib.chol_potrf_f32(A3.data_ptr(), L.data_ptr(), batch, n, n * n, max(n, 1), m,
                                info.data_ptr() if info is not None else None,
                                ws.data_ptr() if ws is not None else None, ws_n, _raw_stream(dev))
    if rc:
        _lib.check(rc, "chol_potrf_f32")
    if return_info:
        return L, info
    return L


def _mixed_with_fallback(A, return_info, out, mode="mixed"):
    global MIXED_FALLBACKS
    L, info = cholesky(A, mode=mode, return_info=True, out=out)
    bad = info.reshape(-1) != 0
    if bad.any().item():
        n = A.shape[-1]
        idx = bad.nonzero().squeeze(1)
        sub = A.reshape(-1, n, n)[idx]
        Lp, ip = cholesky(sub, mode="plain", return_info=True)
        L.view(-1, n, n)[idx] = Lp
        info.view(-1)[idx] = ip
        MIXED_FALLBACKS += int(idx.numel())
    return (L, info) if return_info else L


def custom_kernel(data: torch.Tensor) -> torch.Tensor:
    """Drop-in for the reference's ``custom_kernel`` (submission.py:4101)."""
    return cholesky(data)
