"""Pins the CPU oracle (oracle/) before anything is checked against it.

Each oracle function is checked against something independent of its own
formula: LAPACK (numpy.linalg.cholesky), closed forms, brute force, and
planted mistakes that a plausible bug would make.
"""
import numpy as np
import pytest

RNG = np.random.default_rng(1234)


def _spd(b, n, damping=1.0, seed=0):
    g = np.random.default_rng(seed).standard_normal((b, n, n)).astype(np.float32)
    a = (g @ np.swapaxes(g, -1, -2)) / max(n, 1)
    a += damping * np.eye(n, dtype=np.float32)
    return (0.5 * (a + np.swapaxes(a, -1, -2))).astype(np.float32)


# --------------------------------------------------------------- potrf64 ----
@pytest.mark.parametrize("n", [1, 2, 3, 17, 32, 64, 100, 129])
def test_potrf64_matches_lapack(oracle_mod, n):
    a = _spd(3, n, seed=n)
    L, info = oracle_mod.potrf64(a)
    ref = np.linalg.cholesky(a.astype(np.float64))
    assert (info == 0).all()
    assert np.max(np.abs(L - ref)) <= 1e-12 * max(1.0, np.max(np.abs(ref)))
    assert (np.triu(L, 1) == 0).all()


def test_potrf64_closed_forms(oracle_mod):
    # identity -> identity; diagonal -> sqrt(diagonal); n = 1 -> sqrt(a)
    eye = np.eye(9, dtype=np.float32)[None]
    L, _ = oracle_mod.potrf64(eye)
    assert np.array_equal(L[0], np.eye(9))
    d = np.logspace(0, -5, 16).astype(np.float32)
    L, _ = oracle_mod.potrf64(np.diag(d)[None])
    assert np.allclose(np.diag(L[0]), np.sqrt(d.astype(np.float64)), rtol=1e-15, atol=0)
    assert np.count_nonzero(L[0] - np.diag(np.diag(L[0]))) == 0
    L, _ = oracle_mod.potrf64(np.array([[[4.0]]], np.float32))
    assert L[0, 0, 0] == 2.0


def test_potrf64_tridiagonal_recurrence(oracle_mod):
    # A tridiagonal SPD matrix has a lower-bidiagonal factor given by the
    # O(n) recurrence l_ii = sqrt(d_i - l_{i,i-1}^2), l_{i+1,i} = e_i / l_ii.
    n = 40
    d = RNG.uniform(1.5, 2.5, n).astype(np.float32)
    e = RNG.uniform(-0.25, 0.25, n - 1).astype(np.float32)
    a = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    L, _ = oracle_mod.potrf64(a[None].astype(np.float32))
    ld = np.zeros(n)
    lo = np.zeros(n - 1)
    ld[0] = np.sqrt(np.float64(d[0]))
    for i in range(1, n):
        lo[i - 1] = np.float64(e[i - 1]) / ld[i - 1]
        ld[i] = np.sqrt(np.float64(d[i]) - lo[i - 1] ** 2)
    assert np.allclose(np.diag(L[0]), ld, rtol=1e-14)
    assert np.allclose(np.diag(L[0], -1), lo, rtol=1e-13, atol=1e-16)
    assert np.count_nonzero(np.tril(L[0], -2)) == 0


def test_potrf64_recovers_chosen_factor(oracle_mod):
    # A = L0 L0^T with a chosen well-conditioned L0 (computed exactly in FP64
    # from FP32-representable entries, then rounded to FP32) -> L close to L0.
    n = 48
    L0 = np.tril(RNG.uniform(-0.3, 0.3, (n, n))) + np.diag(RNG.uniform(1, 2, n))
    L0 = L0.astype(np.float32).astype(np.float64)
    a = (L0 @ L0.T).astype(np.float32)
    L, _ = oracle_mod.potrf64(a[None])
    assert np.max(np.abs(L[0] - L0)) < 1e-5


def test_potrf64_reads_lower_triangle_only(oracle_mod):
    a = _spd(2, 20, seed=5)
    b = a.copy()
    iu = np.triu_indices(20, 1)
    b[:, iu[0], iu[1]] = 1e30  # garbage above the diagonal
    La, _ = oracle_mod.potrf64(a)
    Lb, _ = oracle_mod.potrf64(b)
    assert np.array_equal(La, Lb)


@pytest.mark.parametrize("k", [0, 5, 19])
def test_potrf64_info_first_failing_minor(oracle_mod, k):
    a = _spd(1, 20, seed=7)
    a[0, k, k] = -1.0
    L, info = oracle_mod.potrf64(a)
    assert info[0] == k + 1
    assert np.isnan(L[0, k, k])
    if k > 0:
        assert np.isfinite(L[0, :k, :k]).all()


def test_potrf32_close_to_potrf64(oracle_mod):
    a = _spd(2, 96, seed=3)
    L32, _ = oracle_mod.potrf32(a)
    L64, _ = oracle_mod.potrf64(a)
    err = np.max(np.abs(L32 - L64)) / np.max(np.abs(L64))
    assert err < 1e-5
    assert err > 0  # it really is FP32 arithmetic


# -------------------------------------------------------------- residual ----
def test_residual_against_numpy(oracle_mod):
    a = _spd(2, 33, seed=11)
    L = np.tril(RNG.standard_normal((2, 33, 33))).astype(np.float32)
    fro, l1r, l1a = oracle_mod.residual(a, L)
    L64 = L.astype(np.float64)
    R = L64 @ np.swapaxes(L64, -1, -2) - a.astype(np.float64)
    assert np.allclose(fro, np.linalg.norm(R, axis=(1, 2)) / np.linalg.norm(a.astype(np.float64), axis=(1, 2)),
                       rtol=1e-12)
    assert np.allclose(l1r, np.linalg.norm(R, ord=1, axis=(1, 2)), rtol=1e-12)
    assert np.allclose(l1a, np.linalg.norm(a.astype(np.float64), ord=1, axis=(1, 2)), rtol=1e-12)


def test_residual_catches_transposed_factor(oracle_mod):
    a = _spd(1, 24, seed=2)
    L64, _ = oracle_mod.potrf64(a)
    good = L64.astype(np.float32)
    bad = np.swapaxes(good, -1, -2).copy()  # U instead of L: U U^T != A
    assert oracle_mod.residual(a, good)[0][0] < 1e-6
    assert oracle_mod.residual(a, bad)[0][0] > 1e-2


# ----------------------------------------------------------- tri inverse ----
def test_tri_inverse(oracle_mod):
    a = _spd(2, 50, seed=9)
    L, _ = oracle_mod.potrf64(a)
    X = oracle_mod.tri_inverse64(L)
    eye = np.eye(50)
    assert np.max(np.abs(L @ X - eye)) < 1e-12
    assert (np.triu(X, 1) == 0).all()


# --------------------------------------------------------------- checker ----
def test_checker_accepts_exact_factor(oracle_mod):
    a = _spd(4, 40, seed=1)
    L, _ = oracle_mod.potrf64(a)
    ok, msg = oracle_mod.check_implementation(a, L.astype(np.float32))
    assert ok, msg
    assert msg.startswith("scaled_reconstruction_residual=")


def test_checker_rejections(oracle_mod):
    a = _spd(2, 16, seed=4)
    L = oracle_mod.potrf64(a)[0].astype(np.float32)
    ok, msg = oracle_mod.check_implementation(a, L[:1])
    assert not ok and "shape" in msg
    ok, msg = oracle_mod.check_implementation(a, L.astype(np.float64))
    assert not ok and "dtype" in msg
    bad = L.copy(); bad[1, 3, 2] = np.nan
    ok, msg = oracle_mod.check_implementation(a, bad)
    assert not ok and "NaN" in msg
    bad = L.copy(); bad[0, 0, 5] = 0.5  # upper-triangle garbage
    ok, msg = oracle_mod.check_implementation(a, bad)
    assert not ok and "lower triangular" in msg
    bad = L.copy(); bad[1] = -bad[1]  # -L reconstructs A but diag < 0
    ok, msg = oracle_mod.check_implementation(a, bad)
    assert not ok and "diagonal" in msg
    bad = L.copy(); bad[0, 7, 3] += 1e-2  # wrong entry
    ok, msg = oracle_mod.check_implementation(a, bad)
    assert not ok and "reconstruct" in msg


# ------------------------------------------------- planted-mistake oracle ----
def _py_cholesky(a, mistake=None):
    """Pure-Python FP64 Cholesky with an optional planted mistake (small n)."""
    n = a.shape[0]
    L = np.zeros((n, n))
    for i in range(n):
        for j in range(i + 1):
            s = float(a[i, j])
            kmax = j - 1 if mistake == "drop_term" and j > 0 else j
            for k in range(kmax):
                s -= L[i, k] * L[j, k] if mistake != "sign" else -L[i, k] * L[j, k]
            if i == j:
                L[i, i] = np.sqrt(s)
            else:
                L[i, j] = s / (L[i, i] if mistake == "index" else L[j, j])
    return L


@pytest.mark.parametrize("mistake", [None, "drop_term", "sign", "index"])
def test_checks_catch_planted_mistakes(oracle_mod, mistake):
    a = _spd(1, 12, seed=21)[0]
    with np.errstate(invalid="ignore", divide="ignore"):
        L = _py_cholesky(a, mistake)
    ref = np.linalg.cholesky(a.astype(np.float64))
    matches_lapack = np.allclose(L, ref, rtol=1e-10, atol=1e-12)
    fro = oracle_mod.residual(a[None], np.nan_to_num(L, nan=1e3).astype(np.float32)[None])[0][0]
    if mistake is None:
        assert matches_lapack and fro < 1e-6
    else:
        assert not matches_lapack
        assert not fro < 1e-6
