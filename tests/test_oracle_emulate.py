"""Pins the oracle's precision emulations (oracle/emulate.py) and its FP32
recurrence (oracle.potrf32) before any design decision rests on them.

Each function is checked against something its own formula does not
contain: torch's BF16 / FP8-E4M3 casts, exact rational arithmetic, a
per-tile brute force, the FP64 oracle, and planted variants (a dropped
split term) that a plausible mistake would produce.
"""
from fractions import Fraction

import numpy as np
import pytest
import torch

from oracle import emulate

F32 = np.float32


def _probe_values(rng, count=200_000):
    """Random magnitudes over the FP32 exponent range plus the cases rounding
    code gets wrong: exact ties, subnormals, zeros, powers of two."""
    mag = np.exp2(rng.uniform(-140, 120, count)) * rng.choice([-1.0, 1.0], count)
    x = mag.astype(F32)
    # BF16 ties: low 16 bits exactly 0x8000 (odd and even kept halves)
    u = rng.integers(0x00800000, 0x7F000000, 4096, dtype=np.uint32)
    ties = ((u & 0xFFFF0000) | 0x8000).view(F32)
    sub = (rng.integers(1, 0x007FFFFF, 4096, dtype=np.uint32)).view(F32)  # FP32 subnormals
    extra = np.array([0.0, -0.0, 1.0, -1.0, 2.0 ** -126, 2.0 ** 100, 3.0, 1.0 + 2.0 ** -8,
                      1.0 + 2.0 ** -9, 1.0 + 3 * 2.0 ** -9], dtype=F32)
    return np.concatenate([x, ties, -ties, sub, extra])


# --------------------------------------------------------------- BF16 ----
def test_bf16_round_matches_torch_cast():
    x = _probe_values(np.random.default_rng(1))
    ours = emulate.bf16_round(x)
    ref = torch.from_numpy(x).to(torch.bfloat16).float().numpy()
    assert np.array_equal(ours.view(np.uint32) & 0x7FFFFFFF, ref.view(np.uint32) & 0x7FFFFFFF)
    assert np.array_equal(ours, ref)


def test_split_bf16_matches_torch_and_bounds_remainder():
    x = np.random.default_rng(2).standard_normal(100_000).astype(F32) * F32(3.7)
    hi, lo = emulate.split_bf16(x)
    t = torch.from_numpy(x)
    thi = t.to(torch.bfloat16).float()
    tlo = (t - thi).to(torch.bfloat16).float()
    assert np.array_equal(hi, thi.numpy()) and np.array_equal(lo, tlo.numpy())
    # x = hi + lo + r with |r| <= 2^-16 |x| (two BF16 roundings, 8-bit significands)
    r = x.astype(np.float64) - hi.astype(np.float64) - lo.astype(np.float64)
    assert np.all(np.abs(r) <= 2.0 ** -16 * np.abs(x.astype(np.float64)))


def _bf16x3_bound(a, b):
    """Elementwise bound on |bf16x3(a, b) - a b^T| (FP64 exact product): the
    dropped lo*lo term and the two split remainders, each <= 2^-16 |a||b|
    termwise, plus FP32 accumulation of three K-term sums."""
    k = a.shape[1]
    ab = np.abs(a.astype(np.float64)) @ np.abs(b.astype(np.float64)).T
    return (3 * 2.0 ** -16 + 2.0 ** -16 + 4 * k * 2.0 ** -24) * ab


def test_bf16x3_nt_within_split_bound():
    rng = np.random.default_rng(3)
    a = rng.standard_normal((96, 128)).astype(F32)
    b = rng.standard_normal((80, 128)).astype(F32)
    exact = a.astype(np.float64) @ b.astype(np.float64).T
    got = emulate.bf16x3_nt(a, b).astype(np.float64)
    bound = _bf16x3_bound(a, b)
    assert np.all(np.abs(got - exact) <= bound)
    # and it really is better than one BF16 product by orders of magnitude
    h = lambda m: torch.from_numpy(m).to(torch.bfloat16).double().numpy()
    one = h(a) @ h(b).T
    assert np.abs(one - exact).max() > 50 * np.abs(got - exact).max()


def test_bf16x3_nt_planted_dropped_term_fails_bound():
    # A split product that keeps lo*lo but drops hi*lo (a plausible index
    # mistake in the three-term sum) must violate the bound.
    rng = np.random.default_rng(4)
    a = rng.standard_normal((64, 64)).astype(F32)
    b = rng.standard_normal((64, 64)).astype(F32)
    ah, al = emulate.split_bf16(a)
    bh, bl = emulate.split_bf16(b)
    planted = (ah @ bh.T).astype(F32) + (al @ bh.T).astype(F32) + (al @ bl.T).astype(F32)
    exact = a.astype(np.float64) @ b.astype(np.float64).T
    assert np.any(np.abs(planted.astype(np.float64) - exact) > _bf16x3_bound(a, b))


# ------------------------------------------------------------ FP8 E4M3 ----
def test_e4m3_round_matches_torch_cast_in_range():
    # torch.float8_e4m3fn: nearest-even, subnormal step 2^-9, max 448; inputs
    # at or beyond the 464 rounding boundary give NaN there, while the
    # emulation saturates to 448 (tested separately below).
    rng = np.random.default_rng(5)
    x = np.concatenate([
        rng.uniform(-463.0, 463.0, 100_000),
        np.exp2(rng.uniform(-14, 8.85, 100_000)) * rng.choice([-1.0, 1.0], 100_000),
    ]).astype(F32)
    # exact midpoints between neighbouring E4M3 values (ties) and subnormals
    grid = np.unique(torch.arange(0, 127, dtype=torch.uint8).view(torch.float8_e4m3fn).float().numpy())
    mids = ((grid[:-1].astype(np.float64) + grid[1:]) / 2).astype(F32)
    sub = np.arange(0, 16, dtype=np.float64) * 2.0 ** -10
    x = np.concatenate([x, grid, -grid, mids, -mids, sub.astype(F32)])
    x = x[np.abs(x) < 464.0]
    ours = emulate.e4m3_round(x)
    ref = torch.from_numpy(x).to(torch.float8_e4m3fn).float().numpy()
    assert np.array_equal(ours, ref), x[ours != ref][:8]


def test_e4m3_round_saturates():
    x = np.array([448.0, 463.0, 464.0, 465.0, 1e6, -1e6], dtype=F32)
    assert np.array_equal(emulate.e4m3_round(x), np.array([448, 448, 448, 448, 448, -448], dtype=F32))


def test_e4m3_tile_nt_matches_per_tile_brute_force():
    # 2 x 2 tiles of 4 x 4 per operand, each tile with a very different
    # magnitude so a wrong tile index or a shared scale changes the result.
    rng = np.random.default_rng(6)
    a = rng.standard_normal((8, 8)).astype(F32)
    b = rng.standard_normal((8, 8)).astype(F32)
    a[:4, 4:] *= 1000.0
    a[4:, :4] *= 1e-3
    b[4:, 4:] *= 50.0
    got = emulate.e4m3_tile_nt(a, b, tile=4)

    def quant(m):
        out = np.zeros((8, 8), np.float64)
        for ti in range(2):
            for tj in range(2):
                amax = max(abs(float(m[i, j])) for i in range(4 * ti, 4 * ti + 4) for j in range(4 * tj, 4 * tj + 4))
                s = 448.0 / amax
                for i in range(4 * ti, 4 * ti + 4):
                    for j in range(4 * tj, 4 * tj + 4):
                        q = torch.tensor([float(m[i, j]) * s], dtype=torch.float32).to(torch.float8_e4m3fn)
                        out[i, j] = float(q.float()) / s
        return out

    qa, qb = quant(a), quant(b)
    ref = np.array([[sum(qa[i, k] * qb[j, k] for k in range(8)) for j in range(8)] for i in range(8)])
    assert np.allclose(got, ref, rtol=1e-6, atol=1e-6 * np.abs(ref).max())
    # one scale for the whole operand (a planted mistake) gives another answer
    assert not np.allclose(emulate.e4m3_tile_nt(a, b, tile=8), ref, rtol=1e-3, atol=0)


# --------------------------------------------------- blocked_cholesky ----
def _spd(n, seed, damping=0.1):
    g = np.random.default_rng(seed).standard_normal((n, n)).astype(F32)
    a = (g @ g.T) / F32(n) + F32(damping) * np.eye(n, dtype=F32)
    return (0.5 * (a + a.T)).astype(F32)


@pytest.mark.parametrize("n,nb", [(256, 64), (200, 64), (130, 32)])
def test_blocked_cholesky_fp32_matches_fp64_oracle(oracle_mod, n, nb):
    a = _spd(n, 10 + n)
    L = emulate.blocked_cholesky(a, nb=nb, product="fp32")
    L64, info = oracle_mod.potrf64(a[None])
    assert info[0] == 0
    err = np.abs(L - L64[0]).max() / np.abs(L64[0]).max()
    fro = oracle_mod.residual(a[None], L.astype(F32)[None])[0][0]
    assert err < 1e-4 and fro < 8 * np.sqrt(n) * 2.0 ** -23
    assert (np.triu(L, 1) == 0).all()


def test_blocked_cholesky_precision_ladder(oracle_mod):
    # The residuals the mixed-form design rests on, at n = 512 (cond ~1e2):
    # FP32 and BF16x3 products stay below 1e-5, one BF16 product is ~1e-4 or
    # worse, per-tile E4M3 is ~1e-2.  A planted product with the hi*lo term
    # dropped lands at the single-BF16 level, not at BF16x3's.
    n = 512
    a = _spd(n, 77)
    fro = {}
    for prod in ("fp32", "bf16x3", "bf16", "e4m3"):
        L = emulate.blocked_cholesky(a, nb=128, product=prod)
        fro[prod] = oracle_mod.residual(a[None], L.astype(F32)[None])[0][0]
    assert fro["fp32"] < 1e-6
    assert fro["bf16x3"] < 1e-5
    assert fro["bf16"] > 1e-4
    assert fro["e4m3"] > 1e-3
    assert fro["fp32"] <= fro["bf16x3"] < fro["bf16"] < fro["e4m3"]


# ------------------------------------------------------------- potrf32 ----
def _round_f32(x: Fraction) -> np.float32:
    """Exact rational -> nearest FP32, ties to even (no double rounding)."""
    c = np.float32(float(x))
    best = None
    for cand in (np.nextafter(c, np.float32(-np.inf)), c, np.nextafter(c, np.float32(np.inf))):
        d = abs(Fraction(float(cand)) - x)
        key = (d, int(np.array(cand, F32).view(np.uint32)) & 1)
        if best is None or key < best[0]:
            best = (key, cand)
    return np.float32(best[1])


def _fma32(a, b, c) -> np.float32:
    return _round_f32(Fraction(float(a)) * Fraction(float(b)) + Fraction(float(c)))


def _potrf32_steps(a):
    """Row-by-row FP32 recurrence, one correctly rounded operation per step:
    s = fma(-L[i,k], L[j,k], s) for k < j, then sqrt(s) or s / L[j,j]."""
    n = a.shape[0]
    L = np.zeros((n, n), F32)
    for i in range(n):
        for j in range(i + 1):
            s = F32(a[i, j])
            for k in range(j):
                s = _fma32(-L[i, k], L[j, k], s)
            L[i, j] = np.sqrt(s) if i == j else F32(s) / L[j, j]
    return L


@pytest.mark.parametrize("n", [1, 2, 3, 5, 8])
def test_potrf32_bit_exact_against_fp32_steps(oracle_mod, n):
    a = _spd(n, 300 + n, damping=0.05)
    a[np.triu_indices(n, 1)] = 0  # only the lower triangle is read
    L32, info = oracle_mod.potrf32(a[None])
    assert info[0] == 0
    assert np.array_equal(L32[0].view(np.uint32), _potrf32_steps(a).view(np.uint32))


def test_potrf32_is_not_rounded_fp64(oracle_mod):
    # An FP64 recurrence rounded to FP32 at the end (the mistake a
    # closeness-only test cannot see) differs bitwise from the FP32 steps.
    a = _spd(8, 42, damping=0.01)
    L32, _ = oracle_mod.potrf32(a[None])
    L64, _ = oracle_mod.potrf64(a[None])
    assert np.array_equal(L32[0], _potrf32_steps(a))
    assert not np.array_equal(L32[0], L64[0].astype(F32))


# ----------------------------------------------------- correction_step ----
def _perturbed_factor(n, eps, seed):
    rng = np.random.default_rng(seed)
    a = _spd(n, seed, damping=0.5).astype(np.float64)
    L0 = np.linalg.cholesky(a)
    X0 = np.tril(rng.standard_normal((n, n)))
    return a, L0, L0 @ (np.eye(n) + eps * X0)


def test_correction_step_converges_quadratically():
    # From L = L0 (I + eps X0) the corrected factor is within O(eps^2) of the
    # exact factor L0: shrinking eps 10x shrinks the error ~100x.
    errs = []
    for eps in (1e-3, 1e-4):
        a, L0, L = _perturbed_factor(64, eps, 9)
        L1 = emulate.correction_step(a, L)
        errs.append(np.abs(L1 - L0).max() / np.abs(L0).max())
    assert errs[0] < 1e-4 and 50 < errs[0] / errs[1] < 200
    a, L0, _ = _perturbed_factor(64, 0.0, 9)
    assert np.abs(emulate.correction_step(a, L0) - L0).max() < 1e-13  # fixed point


def test_correction_step_planted_full_diagonal_is_first_order():
    # Phi without the halved diagonal (a plausible mistake) over-corrects the
    # diagonal: the error then shrinks only ~10x per 10x in eps.
    errs = []
    for eps in (1e-3, 1e-4):
        a, L0, L = _perturbed_factor(64, eps, 11)
        E = a - L @ L.T
        M = np.linalg.solve(L, np.linalg.solve(L, E).T).T
        L1 = L + L @ np.tril(M)
        errs.append(np.abs(L1 - L0).max() / np.abs(L0).max())
    assert errs[0] / errs[1] < 20


def test_one_correction_step_leaves_fp8_factor_above_target(oracle_mod):
    # The FP8 large-n form: per-128-tile E4M3 trailing products, then
    # one FP64 correction step.  At the large-n workload's input class
    # (G G^T/n + 0.1 I, cond ~40) the E4M3 factor sits near 1e-2 and one
    # step brings it only to ~5e-4, 40x above the 1e-5 residual target; two
    # steps reach ~1e-6.
    n = 512
    a = _spd(n, 517, damping=0.1)
    L = emulate.blocked_cholesky(a, nb=128, product="e4m3")
    r = [oracle_mod.residual(a[None], L.astype(F32)[None])[0][0]]
    for _ in range(2):
        L = emulate.correction_step(a, L)
        r.append(oracle_mod.residual(a[None], L.astype(F32)[None])[0][0])
    assert r[0] > 5e-3
    assert 1e-4 < r[1] < 2e-3
    assert r[2] < 1e-5
