"""numpy emulations of the reduced-precision products of the mixed form.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

* ``bf16_round``: FP32 -> nearest-even BF16, returned as FP32 values.
* ``split_bf16``: x = hi + lo (+ residual ~2^-16 |x|), both BF16-valued.
* ``bf16x3_nt``: A B^T as hiA hiB^T + hiA loB^T + loA hiB^T, products of
  BF16 values (exact in FP32) accumulated in FP32, like the matrix cores.
  This is the 2-term split of scripts/_gpu_runner.py:1684-1707 with the
  lo*lo term dropped (that runner keeps 3 terms and 9 products).
* ``e4m3_round``: FP32 -> OCP FP8 E4M3 (3 mantissa bits, max 448, no inf),
  nearest-even, per the OCP 8-bit floating point specification.
* ``blocked_cholesky``: right-looking blocked Cholesky in FP32 whose
  trailing update uses a chosen product (reference pattern
  scripts/_gpu_runner.py:1710-1762), to measure what each precision does to
  the reconstruction residual before any GPU time is spent.
* ``correction_step``: one first-order refinement of a factor in FP64,
  L + L Phi(L^-1 (A - L L^T) L^-T), the step a low-precision factor would
  need to win its accuracy back.

Each function is pinned in tests/test_oracle_emulate.py (torch's BF16 and
FP8 casts, exact rational arithmetic, per-tile brute force, the FP64
oracle, quadratic convergence, planted mistakes).
"""
from __future__ import annotations

import numpy as np


def bf16_round(x: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    # round to nearest even on the low 16 bits (finite inputs)
    bias = 0x7FFF + ((u >> 16) & 1)
    r = ((u + bias) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).reshape(x.shape)


def split_bf16(x: np.ndarray):
    x = np.asarray(x, dtype=np.float32)
    hi = bf16_round(x)
    lo = bf16_round(x - hi)
    return hi, lo


def bf16x3_nt(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    ah, al = split_bf16(a)
    bh, bl = split_bf16(b)
    f = np.float32
    return (ah @ bh.T).astype(f) + (ah @ bl.T).astype(f) + (al @ bh.T).astype(f)


def e4m3_round(x: np.ndarray) -> np.ndarray:
    """Nearest-even OCP E4M3 (bias 7, 3 mantissa bits, max normal 448,
    min normal 2^-6, subnormal step 2^-9); |x| > 448 saturates to 448."""
    x = np.asarray(x, dtype=np.float64)
    s = np.sign(x)
    a = np.minimum(np.abs(x), 448.0)
    e = np.floor(np.log2(np.maximum(a, 2.0 ** -9)))
    e = np.maximum(e, -6.0)  # subnormals share the 2^-6 exponent
    q = 2.0 ** (e - 3)       # spacing at that exponent
    r = np.round(a / q) * q  # numpy rounds half to even
    return (s * np.minimum(r, 448.0)).astype(np.float32)


def e4m3_tile_nt(a: np.ndarray, b: np.ndarray, tile: int = 128) -> np.ndarray:
    """A B^T with both operands quantised to E4M3 per tile x tile block
    (scale = 448 / absmax of the block), products accumulated in FP32."""
    def q(m):
        out = np.empty_like(m, dtype=np.float32)
        for i in range(0, m.shape[0], tile):
            for j in range(0, m.shape[1], tile):
                blk = m[i:i + tile, j:j + tile]
                amax = float(np.abs(blk).max())
                s = 448.0 / amax if amax > 0 else 1.0
                out[i:i + tile, j:j + tile] = e4m3_round(blk * s) / s
        return out
    return (q(a) @ q(b).T).astype(np.float32)


def blocked_cholesky(a: np.ndarray, nb: int = 64, product: str = "fp32") -> np.ndarray:
    """Right-looking blocked Cholesky of one FP32 matrix; the diagonal block
    and panel in FP32 (panel by the explicit block inverse, as the GPU path),
    the trailing update by ``product`` in {"fp32", "bf16x3", "bf16", "e4m3"}."""
    A = np.tril(np.asarray(a, dtype=np.float32)).copy()
    A = A + np.tril(A, -1).T
    n = A.shape[0]
    f = np.float32
    for k in range(0, n, nb):
        kb = min(nb, n - k)
        L11 = np.linalg.cholesky(A[k:k + kb, k:k + kb].astype(np.float64)).astype(f)
        A[k:k + kb, k:k + kb] = L11
        if k + kb >= n:
            break
        W = np.linalg.inv(L11.astype(np.float64)).astype(f)
        P = (A[k + kb:, k:k + kb] @ W.T).astype(f)
        A[k + kb:, k:k + kb] = P
        if product == "fp32":
            upd = (P @ P.T).astype(f)
        elif product == "bf16x3":
            upd = bf16x3_nt(P, P)
        elif product == "bf16":
            h = bf16_round(P)
            upd = (h @ h.T).astype(f)
        elif product == "e4m3":
            upd = e4m3_tile_nt(P, P)
        else:
            raise ValueError(product)
        A[k + kb:, k + kb:] -= upd
    return np.tril(A)


def phi(m: np.ndarray) -> np.ndarray:
    """Lower triangle of ``m`` with its diagonal halved."""
    p = np.tril(m)
    idx = np.arange(m.shape[-1])
    p[..., idx, idx] *= 0.5
    return p


def correction_step(a: np.ndarray, L: np.ndarray) -> np.ndarray:
    """One first-order correction of a Cholesky factor, all in FP64:
    with M = L^-1 (A - L L^T) L^-T, (I + X)(I + X)^T = I + M is solved to
    first order by the lower-triangular X = Phi(M), so L <- L + L Phi(M).
    The remaining error is second order in M."""
    A = np.asarray(a, dtype=np.float64)
    A = np.tril(A) + np.tril(A, -1).T
    L = np.asarray(L, dtype=np.float64)
    E = A - L @ L.T
    Y = np.linalg.solve(L, E)          # L^-1 E
    M = np.linalg.solve(L, Y.T).T      # (L^-1 E) L^-T
    return L + L @ phi(0.5 * (M + M.T))
