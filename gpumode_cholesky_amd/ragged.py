"""Ragged batches: many matrices of different orders in one call.

``cholesky_ragged(mats)`` factors a list of FP32 SPD device tensors of mixed
orders through ``chol_potrf_ragged_f32``: the library sorts them by order and
runs each size class as one launch sequence (persistent work-queue kernels
for n <= 64, the blocked path in lock step with outer blocking above).  The reference has no
ragged entry point (it factors one fixed (batch, n) per call,
submission.py:4101-4268).
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib
from .cholesky import _MODES, _raw_stream, _validate


def cholesky_ragged(mats, mode: str = "plain", return_info: bool = False, padded: bool = True):
    """Factor each tensor of ``mats`` ([n_i, n_i] FP32, one device).

    Returns the list of factors (and an int32 ``info`` tensor, one entry per
    matrix, when ``return_info``).  The factors are views of one device
    buffer, each at a 256-byte aligned offset: writing one leaves the others
    unchanged, and the buffer is freed when the last of them is.  With
    ``padded`` (the default) the rows of a factor with n > 64 are padded to
    a multiple of 4 floats (a strided view, stride (ld, 1)), which gives the
    kernels float4 access whatever the order; ``padded=False`` returns
    contiguous factors."""
    m = _MODES.get(mode)
    if m is None:
        raise ValueError(f"mode must be one of {sorted(_MODES)}, got {mode!r}")
    if not mats:
        return ([], torch.empty(0, dtype=torch.int32)) if return_info else []
    _validate(mats[0])
    dev = mats[0].device
    # One pass over the list with the cheapest checks (hundreds of matrices:
    # the host time is part of the call); a failing matrix gets the full
    # check for its message.
    f32, idx = torch.float32, dev.index
    ns, ptrs, keep = [], [], []
    for a in mats:
        s = a.shape
        if not isinstance(a, torch.Tensor) or a.dtype is not f32 or len(s) != 2 or s[0] != s[1] \
                or a.get_device() != idx:
            _validate(a)
            raise ValueError("every matrix must be a 2-D [n, n] tensor on one device")
        if not a.is_contiguous():
            a = a.contiguous()
            keep.append(a)  # alive until the kernels have read it (same stream)
        ns.append(s[0])
        ptrs.append(a.data_ptr())
    count = len(mats)
    n_host = np.array(ns, dtype=np.int64)
    ld_host = np.where(n_host > 64, (n_host + 3) // 4 * 4, n_host) if padded else n_host.copy()
    a_host = np.array(ptrs, dtype=np.uint64)
    # One output allocation instead of one per matrix (the host time of
    # hundreds of allocations is a visible share of a ragged call); the
    # views are made after the kernels are enqueued, while they run.
    span = (n_host * ld_host + 63) // 64 * 64
    offs = np.zeros(count, dtype=np.int64)
    np.cumsum(span[:-1], out=offs[1:])
    flat = torch.empty((max(int(span.sum()), 1),), dtype=torch.float32, device=dev)
    l_host = np.uint64(flat.data_ptr()) + 4 * offs.astype(np.uint64)
    lib = _lib.load()
    nws = lib.chol_ragged_workspace_bytes(n_host.ctypes.data, count)
    ws = torch.empty((max(nws, 1),), dtype=torch.uint8, device=dev)
    info = torch.empty((count,), dtype=torch.int32, device=dev) if return_info else None
    with torch.cuda.device(dev):
        rc = lib.chol_potrf_ragged_f32(a_host.ctypes.data, l_host.ctypes.data, n_host.ctypes.data,
                                       ld_host.ctypes.data, count, m,
                                       info.data_ptr() if info is not None else None,
                                       ws.data_ptr(), nws, _raw_stream(dev.index))
    _lib.check(rc, "chol_potrf_ragged_f32")
    outs = [flat.as_strided((n, n), (ld, 1), o) for n, ld, o in zip(n_host.tolist(), ld_host.tolist(), offs.tolist())]
    return (outs, info) if return_info else outs
