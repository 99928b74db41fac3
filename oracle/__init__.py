"""CPU oracle for the batched Cholesky path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
CPU-baseline leg of ``bench.py`` may import this package, and only as the
checker (or, for the baseline, as the timed CPU port).  The product path in
``gpumode_cholesky_amd`` never imports it and has no CPU fallback.

Contents
--------
* ``potrf64``: FP64 Cholesky of FP32 inputs (C, ``cholesky_oracle.c``), the
  ground truth ``ref_kernel = torch.linalg.cholesky_ex(...).L`` of
  reference/reference.py:129-130 restated as the textbook recurrence.
* ``potrf32``: the same recurrence rounded to FP32 at every step.
* ``residual``: FP64 ``||LL^T-A||_F/||A||_F`` and the L1 norms used by the
  reference checker (reference/reference.py:133-134, :170-185).
* ``check_implementation``: numpy restatement of the reference's gate
  (reference/reference.py:137-188).
* ``tri_inverse64``: FP64 inverse of a lower-triangular factor.
* ``emulate``: numpy emulations of split-BF16 and FP8-e4m3 products.

Parity of each function is pinned in tests/test_oracle.py (LAPACK via numpy,
closed forms, brute force, planted mistakes); see the module docstrings.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

# reference/reference.py:6-7
RECON_RTOL_FACTOR = 20.0
TRIANGULAR_RTOL_FACTOR = 8.0
F32_EPS = float(np.finfo(np.float32).eps)
F32_TINY = float(np.finfo(np.float32).tiny)


def build() -> str:
    """Compile the C oracle with its Makefile (gcc, no GPU needed)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    lib = ctypes.CDLL(_LIB_PATH)
    p = ctypes.c_void_p
    i64 = ctypes.c_int64
    lib.oracle_potrf_f64.argtypes = [p, p, i64, i64, p]
    lib.oracle_potrf_f32.argtypes = [p, p, i64, i64, p]
    lib.oracle_residual.argtypes = [p, p, i64, i64, p, p, p]
    lib.oracle_tri_inverse_f64.argtypes = [p, p, i64, i64]
    for f in (lib.oracle_potrf_f64, lib.oracle_potrf_f32, lib.oracle_residual,
              lib.oracle_tri_inverse_f64):
        f.restype = ctypes.c_int
    _lib = lib
    return lib


def _as3(a: np.ndarray, dtype) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=dtype)
    if a.ndim == 2:
        a = a[None]
    if a.ndim != 3 or a.shape[-1] != a.shape[-2]:
        raise ValueError(f"expected [batch, n, n], got {a.shape}")
    return a


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def potrf64(A: np.ndarray):
    """FP64 Cholesky of FP32 matrices (lower triangle read).

    Returns ``(L, info)``: ``L`` float64 [b, n, n] with zeros above the
    diagonal, ``info`` int32 [b] (0, or the 1-based first failing minor).
    """
    A = _as3(A, np.float32)
    b, n, _ = A.shape
    L = np.empty((b, n, n), np.float64)
    info = np.zeros(b, np.int32)
    rc = _load().oracle_potrf_f64(_ptr(A), _ptr(L), b, n, _ptr(info))
    if rc != 0:
        raise RuntimeError(f"oracle_potrf_f64 failed: {rc}")
    return L, info


def potrf32(A: np.ndarray):
    """The same recurrence with every operation in FP32 (fmaf)."""
    A = _as3(A, np.float32)
    b, n, _ = A.shape
    L = np.empty((b, n, n), np.float32)
    info = np.zeros(b, np.int32)
    rc = _load().oracle_potrf_f32(_ptr(A), _ptr(L), b, n, _ptr(info))
    if rc != 0:
        raise RuntimeError(f"oracle_potrf_f32 failed: {rc}")
    return L, info


def residual(A: np.ndarray, L: np.ndarray):
    """FP64 residuals of an FP32 factor: (fro_rel, l1_res, l1_A) per matrix."""
    A = _as3(A, np.float32)
    L = _as3(L, np.float32)
    if A.shape != L.shape:
        raise ValueError(f"shape mismatch {A.shape} vs {L.shape}")
    b, n, _ = A.shape
    fro = np.empty(b, np.float64)
    l1r = np.empty(b, np.float64)
    l1a = np.empty(b, np.float64)
    rc = _load().oracle_residual(_ptr(A), _ptr(L), b, n, _ptr(fro), _ptr(l1r),
                                 _ptr(l1a))
    if rc != 0:
        raise RuntimeError(f"oracle_residual failed: {rc}")
    return fro, l1r, l1a


def tri_inverse64(L: np.ndarray) -> np.ndarray:
    """FP64 inverse of lower-triangular matrices."""
    L = _as3(L, np.float64)
    b, n, _ = L.shape
    X = np.empty((b, n, n), np.float64)
    rc = _load().oracle_tri_inverse_f64(_ptr(L), _ptr(X), b, n)
    if rc != 0:
        raise RuntimeError(f"oracle_tri_inverse_f64 failed: {rc}")
    return X


def _l1(m: np.ndarray) -> np.ndarray:
    """Matrix 1-norm (max column abs-sum), reference/reference.py:133-134."""
    return np.abs(m).sum(axis=-2).max(axis=-1)


def check_implementation(data: np.ndarray, output: np.ndarray):
    """numpy restatement of reference/reference.py:137-188 (FP32 arithmetic).

    Returns ``(ok, message)`` with the same messages as the reference.
    """
    data = np.asarray(data)
    output = np.asarray(output)
    if output.shape != data.shape:  # :140-144
        return False, (f"output shape must be {tuple(data.shape)}, "
                       f"got {tuple(output.shape)}")
    if output.dtype != np.float32:  # :145-146
        return False, f"output dtype must be torch.float32, got {output.dtype}"
    if not np.isfinite(output).all():  # :149-150
        return False, "output contains NaN or Inf"
    batch, n, _ = data.shape
    eps = F32_EPS
    scale = np.maximum(_l1(data.astype(np.float32)), np.float32(F32_TINY))
    upper = np.triu(output, k=1)  # :156
    tri_res = _l1(upper)
    tri_allowed = TRIANGULAR_RTOL_FACTOR * max(n, 1) * eps * scale
    if np.any(tri_res > tri_allowed):  # :159-164
        worst = float((tri_res / scale).max())
        return False, ("output is not lower triangular enough: "
                       f"relative_residual={worst:.3g}")
    diag = np.diagonal(output, axis1=-2, axis2=-1)
    if np.any(diag <= 0):  # :166-168
        return False, "output diagonal must be strictly positive"
    recon = np.matmul(output, np.swapaxes(output, -1, -2))  # :176, FP32
    res = _l1(recon - data)
    allowed = RECON_RTOL_FACTOR * max(n, 1) * eps * scale
    if np.any(res > allowed):  # :182-185
        worst = float((res / scale).max())
        return False, ("L @ L.T does not reconstruct A: "
                       f"relative_residual={worst:.3g}")
    scaled = float((res / (eps * max(n, 1) * scale)).max())
    return True, f"scaled_reconstruction_residual={scaled:.3g}; batch={batch}; n={n}"
