"""The harness's restated input families (reference/reference.py:37-126),
pinned by what each family's construction fixes rather than by retyping it:
eigenvalues, closed forms, structure, ranges and seeding."""
import numpy as np
import pytest
import torch

from gpumode_cholesky_amd import harness


def _gen(case, batch, n, cond, seed):
    return harness.generate_input(batch, n, cond, seed, case, device="cpu")


@pytest.mark.parametrize("case", ["dense", "diagonal", "spectrum", "lowrank", "rowscale", "tridiagonal"])
@pytest.mark.parametrize("n", [1, 7, 40])
def test_every_family_symmetric_positive_definite(case, n):
    a = _gen(case, 3, n, 3, 100 + n)
    assert a.shape == (3, n, n) and a.dtype == torch.float32 and a.is_contiguous()
    x = a.double().numpy()
    assert np.array_equal(x, np.swapaxes(x, -1, -2))
    assert np.linalg.eigvalsh(x).min() > 0


@pytest.mark.parametrize("case", ["dense", "spectrum", "lowrank", "rowscale", "tridiagonal"])
def test_seeding(case):
    a = _gen(case, 2, 12, 3, 5)
    assert torch.equal(a, _gen(case, 2, 12, 3, 5))
    assert not torch.equal(a, _gen(case, 2, 12, 3, 6))
    assert not torch.equal(a[0], a[1])  # each matrix of the batch its own draw


def test_diagonal_closed_form():
    a = _gen("diagonal", 2, 5, 4, 1).double().numpy()
    want = np.diag(10.0 ** np.linspace(0.0, -4.0, 5))
    np.testing.assert_allclose(a[1], want, rtol=1e-6, atol=0)


@pytest.mark.parametrize("cond", [1, 3, 5])
def test_spectrum_eigenvalues_are_the_logspace(cond):
    n = 24
    a = _gen("spectrum", 2, n, cond, 9).double().numpy()
    ev = np.sort(np.linalg.eigvalsh(a), axis=-1)[:, ::-1]
    want = 10.0 ** np.linspace(0.0, -float(cond), n)
    np.testing.assert_allclose(ev, np.broadcast_to(want, ev.shape), rtol=0, atol=3e-6)


def test_dense_damping_and_gram_structure():
    # A - d I = G G^T / n is positive semidefinite with rank n for square G;
    # the smallest eigenvalue of A is therefore >= d (cond 2 -> d = 1e-2).
    a = _gen("dense", 2, 16, 2, 3).double().numpy()
    assert np.linalg.eigvalsh(a).min() >= 1e-2 * (1 - 1e-4)
    # the mean diagonal of G G^T / n is ~1 for standard normal G
    assert 0.6 < np.diagonal(a, axis1=-2, axis2=-1).mean() < 1.4


def test_lowrank_rank_plus_shift():
    n, cond = 64, 3
    a = _gen("lowrank", 1, n, cond, 4).double().numpy()[0]
    ev = np.sort(np.linalg.eigvalsh(a))
    rank = max(1, min(n // 8, 64))
    shift = 10.0 ** -cond
    # n - rank eigenvalues equal the shift, the rest lie above it
    np.testing.assert_allclose(ev[: n - rank], shift, rtol=1e-3)
    assert ev[n - rank] > 10 * shift


def test_rowscale_is_a_congruence_of_a_dense_matrix():
    n, cond = 20, 4
    a = _gen("rowscale", 1, n, cond, 8).double().numpy()[0]
    s = 10.0 ** np.linspace(0.0, -0.5 * cond, n)
    base = a / np.outer(s, s)  # undo the row/column scaling (diagonal shift aside)
    base[np.diag_indices(n)] -= 1e-6 / s**2
    assert np.linalg.eigvalsh(base).min() >= 1e-3 * (1 - 1e-2)
    # the scaling spreads the diagonal by ~ s_0^2 / s_{n-1}^2 = 10^cond
    d = np.diag(a)
    assert d[0] / d[-1] > 10 ** (cond - 1)


def test_tridiagonal_structure_and_ranges():
    n = 30
    a = _gen("tridiagonal", 4, n, 2, 2).numpy()
    d = np.diagonal(a, axis1=-2, axis2=-1)
    off = np.diagonal(a, offset=1, axis1=-2, axis2=-1)
    assert (d >= 1.5).all() and (d <= 2.5).all()
    assert (np.abs(off) <= 0.25).all() and np.abs(off).max() > 0
    band = np.zeros_like(a)
    for k in (-1, 0, 1):
        idx = np.arange(max(0, -k), n - max(0, k))
        band[:, idx, idx + k] = a[:, idx, idx + k]
    assert np.array_equal(a, band)


def test_unknown_family_rejected():
    with pytest.raises(ValueError):
        _gen("banded", 1, 4, 2, 1)
