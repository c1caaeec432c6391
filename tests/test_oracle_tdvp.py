"""Pins of the oracle's Krylov exponential and chain TDVP (CPU only).

* krylov_expm against scipy.linalg.expm (real, imaginary and complex time,
  Hermitian operators), including sub-stepped cases and the tol bound.
* tdvp_chain: at full bond rank the projector-splitting integrator is exact
  (Lubich-Oseledets-Vandereycken exactness property), so the result must equal
  expm(-i H t) psi0 of the dense Hamiltonian, for nsites 1 and 2 and orders 1,
  2 and 4; a sign slip in the backward steps or a wrong plan breaks it.
* with truncation: the state stays normalised (unitary evolution) and an
  eigenstate only picks up the phase exp(-i E t).
* runge_kutta_4 has a local error O(t^5): halving t divides it by ~32.
"""
import numpy as np
import pytest
import scipy.linalg as sl

from oracle import chain, models, tdvp


def _herm(n, seed, cplx=False):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, n))
    if cplx:
        A = A + 1j * rng.standard_normal((n, n))
    return (A + A.conj().T) / 2


@pytest.mark.parametrize("tau", [0.3, -0.2, -0.4j, 0.1 - 0.5j])
@pytest.mark.parametrize("cplx", [False, True])
def test_krylov_expm_matches_expm(tau, cplx):
    A = _herm(80, 1, cplx)
    rng = np.random.default_rng(2)
    x = rng.standard_normal(80) + (1j * rng.standard_normal(80) if (cplx or isinstance(tau, complex)) else 0)
    y, info = tdvp.krylov_expm(lambda v: A @ v, tau, x, krylovdim=30, tol=1e-12)
    want = sl.expm(tau * A) @ x
    assert np.linalg.norm(y - want) / np.linalg.norm(want) < 1e-11
    assert info["numops"] >= 1


def test_krylov_expm_substeps_and_tolerance():
    A = _herm(120, 3) * 4.0
    x = np.random.default_rng(4).standard_normal(120).astype(complex)
    y, info = tdvp.krylov_expm(lambda v: A @ v, -3.0j, x, krylovdim=12, maxiter=500, tol=1e-10)
    assert info["substeps"] > 1
    want = sl.expm(-3.0j * A) @ x
    assert np.linalg.norm(y - want) / np.linalg.norm(x) < 1e-8


def test_krylov_expm_invariant_subspace():
    """A rank-2 operator: x, Ax, A^2 x span an invariant subspace, so the basis
    closes after three vectors (a fourth only from round-off) and the result is exact."""
    u = np.random.default_rng(5).standard_normal((50, 2))
    A = u @ u.T
    x = np.random.default_rng(6).standard_normal(50)
    y, info = tdvp.krylov_expm(lambda v: A @ v, 0.7, x)
    assert info["kmax"] <= 4
    np.testing.assert_allclose(y, sl.expm(0.7 * A) @ x, rtol=1e-12, atol=1e-12)


def test_krylov_expm_zero_time_and_maxiter():
    x = np.arange(5.0)
    y, _ = tdvp.krylov_expm(lambda v: v, 0.0, x)
    np.testing.assert_array_equal(y, x)
    A = _herm(60, 7) * 50
    with pytest.raises(RuntimeError):
        tdvp.krylov_expm(lambda v: A @ v, -5j, x.repeat(12).astype(complex), krylovdim=4, maxiter=3)


@pytest.mark.parametrize("nsites", [1, 2])
@pytest.mark.parametrize("order", [1, 2, 4])
def test_tdvp_exact_at_full_rank(nsites, order):
    N = 8
    W = models.xxz_mpo(N, Jz=0.7, hz=0.1)
    mps = chain.random_mps(N, 2, 64, seed=3)
    psi0 = chain.mps_to_dense(mps).reshape(-1)
    H = chain.mpo_to_dense(W)
    ts = [0.0, 0.05, 0.1]
    out, t = tdvp.tdvp_chain(W, mps, ts, nsites=nsites, tdvp_order=order)
    want = sl.expm(-1j * H * 0.1) @ psi0
    assert np.linalg.norm(chain.mps_to_dense(out).reshape(-1) - want) < 1e-11
    # current_time adds every region's step (2-site: net dt per sweep; 1-site: N dt)
    assert abs(t - (-0.1j) * (1 if nsites == 2 else N)) < 1e-12


def test_tdvp_real_exponents_imaginary_time():
    """applyexp with real exponents: exp(-tau H) at full rank equals the dense one."""
    N = 6
    W = models.heisenberg_mpo(N)
    mps = chain.random_mps(N, 2, 8, seed=9)
    H = chain.mpo_to_dense(W)
    psi0 = chain.mps_to_dense(mps).reshape(-1)
    out, _ = tdvp.applyexp_chain(W, mps, [0.0, -0.2], nsites=2, tdvp_order=2)
    want = sl.expm(-0.2 * H) @ psi0
    assert np.linalg.norm(chain.mps_to_dense(out).reshape(-1) - want) / np.linalg.norm(want) < 1e-11


def test_tdvp_truncated_phase_of_eigenstate():
    """The reference regression (test_tree_tdvp.jl:48-77) on a chain: evolving the
    ground state keeps the norm and only rotates the phase by E t."""
    N = 10
    W = models.heisenberg_mpo(N)
    res = chain.dmrg_chain(W, chain.product_mps([0, 1] * (N // 2), 2), nsweeps=6, maxdim=[10, 20, 100],
                           cutoff=1e-10)
    gs = res.mps
    E = res.energy
    ts = [0.0, 0.02, 0.04, 0.06, 0.08, 0.10]
    for nsites in (1, 2):
        out, _ = tdvp.tdvp_chain(W, gs, ts, nsites=nsites, maxdim=100, cutoff=1e-10)
        nrm = np.sqrt(abs(chain.mps_inner(out, out)))
        assert nrm > 0.999
        z = chain.mps_inner(out, gs)
        assert abs(z) > 0.99
        assert abs(np.angle(z) - E * 0.1) < 1e-4


def test_runge_kutta_4_order():
    A = _herm(30, 11)
    x = np.random.default_rng(12).standard_normal(30).astype(complex)
    errs = []
    for t in (-0.02j, -0.01j):
        y = tdvp.runge_kutta_4(lambda v: A @ v, t, x)
        errs.append(np.linalg.norm(y - sl.expm(t * A) @ x))
    assert 20 < errs[0] / errs[1] < 40
