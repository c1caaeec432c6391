"""Pins the oracle's restated linear algebra (oracle/linalg.py) against hand-computed
cases, numpy/scipy routines and invariants -- not against a retyped formula."""
import numpy as np
import pytest
import scipy.linalg as sl

from oracle import linalg


def test_truncate_maxdim_then_relative_cutoff():
    P = np.array([0.5, 0.3, 0.1, 0.06, 0.03, 0.01])   # sum = 1
    # maxdim=4 drops 0.03, 0.01 -> truncerr 0.04; cutoff 0.1: 0.04+0.06=0.10 <= 0.1 -> drop
    n, err, _ = linalg.truncate(P, maxdim=4, cutoff=0.1)
    assert n == 3 and err == pytest.approx(0.10)
    # cutoff slightly below the accumulated weight keeps the 4th value
    n, err, _ = linalg.truncate(P, maxdim=4, cutoff=0.0999)
    assert n == 4 and err == pytest.approx(0.04)


def test_truncate_mindim_floor_and_negative_tail():
    P = np.array([1.0, 1e-20, -1e-18, -1e-17])
    n, err, _ = linalg.truncate(P, mindim=2, cutoff=1e-3)
    assert n == 2            # mindim stops the cutoff loop
    assert err == 0.0        # the two negative tail entries were zeroed before dropping
    n, err, _ = linalg.truncate([2.0], cutoff=0.9)
    assert n == 1 and err == 0.0   # length-1 spectra are never truncated


def test_truncate_zero_cutoff_keeps_everything_positive():
    P = np.array([3.0, 2.0, 1.0])
    assert linalg.truncate(P, cutoff=0.0)[0] == 3


def test_route_choice_threshold():
    assert linalg.choose_route(0.0) == "svd"
    assert linalg.choose_route(1e-12) == "svd"
    assert linalg.choose_route(1.0000001e-12) == "eigen"
    assert linalg.choose_route(1e-10) == "eigen"
    assert linalg.choose_route(1e-10, which_decomp="svd") == "svd"


@pytest.mark.parametrize("route", ["svd", "eigen"])
@pytest.mark.parametrize("shape", [(12, 20), (20, 12), (16, 16)])
def test_factorize_identities(route, shape):
    rng = np.random.default_rng(7)
    # graded spectrum so that cutoff truncation is exercised
    m, n = shape
    r = min(m, n)
    U0, _ = np.linalg.qr(rng.standard_normal((m, r)))
    V0, _ = np.linalg.qr(rng.standard_normal((n, r)))
    s = 10.0 ** (-np.arange(r) / 2.0)
    M = (U0 * s) @ V0.T
    U, C, info = linalg.factorize(M, cutoff=3e-8, maxdim=9, which_decomp=route)
    k = info.kept
    assert U.shape == (m, k) and C.shape == (k, n)
    np.testing.assert_allclose(U.T @ U, np.eye(k), atol=1e-12)
    err = np.linalg.norm(M - U @ C) ** 2 / np.linalg.norm(M) ** 2
    assert err == pytest.approx(info.truncerr, rel=1e-6, abs=1e-15)
    # spectrum equals the squared singular values
    np.testing.assert_allclose(np.sort(info.spectrum)[::-1][:r], s ** 2, rtol=1e-9, atol=1e-15)
    # the kept count follows truncate() on the exact spectrum
    assert k == linalg.truncate(s ** 2, maxdim=9, cutoff=3e-8)[0]


def test_factorize_rank_deficient_and_degenerate():
    rng = np.random.default_rng(3)
    A = rng.standard_normal((10, 3))
    M = A @ rng.standard_normal((3, 8))   # rank 3
    for route in ("svd", "eigen"):
        U, C, info = linalg.factorize(M, cutoff=1e-10, which_decomp=route)
        assert info.kept == 3
        np.testing.assert_allclose(U @ C, M, atol=1e-10)
    M = np.eye(6)                          # fully degenerate spectrum
    U, C, info = linalg.factorize(M, cutoff=1e-10, maxdim=4, which_decomp="svd")
    assert info.kept == 4 and info.truncerr == pytest.approx(2 / 6)


def test_qr_pos_unique():
    rng = np.random.default_rng(0)
    M = rng.standard_normal((9, 5))
    Q, R = linalg.qr_pos(M)
    np.testing.assert_allclose(Q @ R, M, atol=1e-13)
    np.testing.assert_allclose(Q.T @ Q, np.eye(5), atol=1e-13)
    assert np.all(np.diag(R) >= 0) and np.allclose(np.tril(R, -1), 0)
    M = rng.standard_normal((4, 7))
    Q, R = linalg.qr_pos(M)
    assert Q.shape == (4, 4) and R.shape == (4, 7)


def test_lanczos_exact_when_dimension_le_krylovdim():
    rng = np.random.default_rng(1)
    for n in (1, 2, 3):
        A = rng.standard_normal((n, n)); A = A + A.T
        x0 = rng.standard_normal(n)
        e, x, info = linalg.lanczos_eigsolve(lambda v: A @ v, x0, krylovdim=3)
        assert e == pytest.approx(np.linalg.eigvalsh(A)[0], abs=1e-12)
        np.testing.assert_allclose(A @ x, e * x, atol=1e-10)
        assert info.numops == n


def test_lanczos_matches_explicit_krylov_rayleigh_ritz():
    """The K=3 Ritz value equals the lowest eigenvalue of the operator projected on
    span{x0, Ax0, A^2 x0} (computed here with an orthonormal basis from numpy QR)."""
    rng = np.random.default_rng(5)
    n = 40
    A = rng.standard_normal((n, n)); A = A + A.T
    x0 = rng.standard_normal(n)
    e, x, info = linalg.lanczos_eigsolve(lambda v: A @ v, x0, krylovdim=3)
    Kry = np.stack([x0, A @ x0, A @ A @ x0], axis=1)
    Q, _ = np.linalg.qr(Kry)
    ritz = np.linalg.eigvalsh(Q.T @ A @ Q)
    assert e == pytest.approx(ritz[0], rel=1e-12)
    assert np.linalg.norm(x) == pytest.approx(1.0, abs=1e-12)
    assert x @ x0 > 0  # gauge: Ritz vector overlaps the start vector positively
    # tridiagonal coefficients reproduce scipy's tridiagonal eigenvalues
    w = sl.eigh_tridiagonal(info.alphas, info.betas, eigvals_only=True)
    assert w[0] == pytest.approx(e, rel=1e-14)


def test_lanczos_invariant_subspace_stop():
    """x0 inside a 2-dimensional invariant subspace: beta <= tol after 2 vectors."""
    n = 12
    A = np.diag(np.arange(1.0, n + 1))
    x0 = np.zeros(n); x0[0] = 1.0; x0[4] = 2.0
    e, x, info = linalg.lanczos_eigsolve(lambda v: A @ v, x0, krylovdim=3, tol=1e-14)
    assert info.krylov_dim == 2 and info.numops == 2
    assert e == pytest.approx(1.0, abs=1e-13)


def test_tridiag_gauge_positive_first_component():
    w, y = linalg.tridiag_eigh([1.0, 2.0, 3.0], [0.5, 0.25])
    assert np.all(y[0] > 0)
    T = np.diag([1.0, 2.0, 3.0]) + np.diag([0.5, 0.25], 1) + np.diag([0.5, 0.25], -1)
    np.testing.assert_allclose(T @ y, y * w, atol=1e-14)
