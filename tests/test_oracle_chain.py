"""Pins the oracle's chain tensor algebra and DMRG driver (oracle/chain.py) against
brute-force matrices, ED fixtures and closed forms."""
import json
import os

import numpy as np
import pytest

from oracle import chain, ed, models

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "ed_energies.json")))


def _envs(W, mps, lo, hi):
    L = np.ones((1, 1, 1))
    for i in range(lo):
        L = chain.env_left(L, mps[i], W[i])
    R = np.ones((1, 1, 1))
    for i in range(len(W) - 1, hi, -1):
        R = chain.env_right(R, mps[i], W[i])
    return L, R


@pytest.mark.parametrize("dtype", [np.float64, np.complex128])
def test_heff_matches_kronecker_matrix_and_is_hermitian(dtype):
    W = models.xxz_mpo(6, Jz=0.7, hz=0.3)
    mps = chain.random_mps(6, 2, 3, seed=3, dtype=dtype)
    L, R = _envs(W, mps, 2, 3)
    th = chain.theta2(mps[2], mps[3])
    H = chain.heff2_dense_matrix(L, W[2], W[3], R)
    y = chain.heff2_apply(L, W[2], W[3], R, th)
    np.testing.assert_allclose(H @ th.reshape(-1), y.reshape(-1), atol=1e-13)
    np.testing.assert_allclose(H, H.conj().T, atol=1e-13)


def test_environment_split_equals_dense_expectation():
    N = 7
    W = models.xxz_mpo(N, Jz=1.3, hz=-0.2)
    mps = chain.random_mps(N, 2, 4, seed=11)
    v = chain.mps_to_dense(mps)
    e_dense = v @ chain.mpo_to_dense(W) @ v
    for bond in range(N - 1):
        L, R = _envs(W, mps, bond + 1, bond)
        # contract L (sites <= bond) with R (sites > bond) across the bond
        val = np.einsum("xwa,xwa->", L, R)
        assert val == pytest.approx(e_dense, rel=1e-12)
    # one-site and two-site H_eff Rayleigh quotients equal the same number
    L, R = _envs(W, mps, 3, 4)
    th = chain.theta2(mps[3], mps[4])
    assert np.vdot(th, chain.heff2_apply(L, W[3], W[4], R, th)) == pytest.approx(e_dense, rel=1e-12)
    L, R = _envs(W, mps, 3, 3)
    assert np.vdot(mps[3], chain.heff1_apply(L, W[3], R, mps[3])) == pytest.approx(e_dense, rel=1e-12)


def test_mpo_dense_matches_sector_hamiltonian_spectrum():
    N = 6
    W = models.xxz_mpo(N, Jz=0.4)
    Hd = chain.mpo_to_dense(W)
    Hs, _ = ed.xxz_hamiltonian(N, ed.chain_edges(N), Jz=0.4, m2_total=0)
    assert np.linalg.eigvalsh(Hd)[0] == pytest.approx(ed.ground_energy(Hs), abs=1e-12)


def test_xx_chain_free_fermion_closed_form():
    """Jz = 0: E0 = sum of the negative single-particle energies cos(k pi/(N+1))."""
    N = 10
    eps = np.cos(np.arange(1, N + 1) * np.pi / (N + 1))
    exact = eps[eps < 0].sum()
    assert GOLDEN["xxz_s12_chain"]["10_0.0"] == pytest.approx(exact, abs=1e-11)
    W = models.xxz_mpo(N, Jz=0.0)
    mps = chain.random_mps(N, 2, 4, seed=1)
    r = chain.dmrg_chain(W, mps, nsweeps=6, maxdim=[16, 32, 64], cutoff=1e-14)
    assert r.energy == pytest.approx(exact, abs=1e-10)


def test_n4_closed_form():
    assert GOLDEN["heisenberg_s12_chain"]["4"] == pytest.approx(-(3 + 2 * np.sqrt(3)) / 4, abs=1e-13)


def test_s1_chain_reference_constant():
    """examples/dmrg.jl:41-43 prints -12.8945601 for the S=1 N=10 chain."""
    assert GOLDEN["heisenberg_s1_chain"]["10"] == pytest.approx(-12.8945601, abs=1e-7)
    W = models.heisenberg_mpo(10, d=3)
    mps = chain.random_mps(10, 3, 4, seed=2)
    r = chain.dmrg_chain(W, mps, nsweeps=5, maxdim=[10, 40, 80, 160], cutoff=1e-12)
    assert r.energy == pytest.approx(-12.8945601, abs=1e-7)


def test_dmrg_c1_heisenberg_n20_vs_ed():
    """The CPU-runnable configuration: N=20, maxdim [10,20,40,64,64], cutoff 1e-10, K=3."""
    W = models.heisenberg_mpo(20)
    mps = chain.random_mps(20, 2, 4, seed=0)
    r = chain.dmrg_chain(W, mps, nsweeps=5, maxdim=[10, 20, 40, 64, 64], cutoff=1e-10)
    Eex = GOLDEN["heisenberg_s12_chain"]["20"]
    assert abs(r.energy - Eex) < 1e-8
    assert r.energy >= Eex - 1e-12          # variational
    assert all(l.route == "eigen" for l in r.log)
    assert len(r.log) == 5 * 2 * 19          # 2(N-1) regions per sweep
    assert r.log[0].region == (19, 18) and r.log[18].region == (1, 0) and r.log[19].region == (0, 1)
    assert max(r.sweep_maxlinkdim) <= 64


def test_random_mps_normalised_and_seeded():
    a = chain.random_mps(8, 2, 5, seed=4)
    b = chain.random_mps(8, 2, 5, seed=4)
    assert abs(chain.mps_inner(a, a) - 1.0) < 1e-12
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    assert [t.shape[2] for t in a] == [2, 4, 5, 5, 5, 4, 2, 1]
