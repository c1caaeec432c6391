"""End-to-end 2-site DMRG on the device vs the oracle and ED.

Configuration "C1": spin-1/2 Heisenberg chain, N=20, open boundaries, MPO bond
dim 5, dense FP64, random MPS of bond dim 4 (seed 0), 5 sweeps,
maxdim [10, 20, 40, 64, 64], cutoff 1e-10, Lanczos krylovdim 3, maxiter 1.
Bar: every region's Ritz value within 1e-10 relative of the oracle's, identical
kept bond dimensions and decomposition routes, truncation errors within 1e-6
relative or 1e-13 absolute (sums of eigenvalues of rho, each known to about
eps * tr(rho) on the eigen route), final energy within 1e-8
of exact diagonalisation (the cutoff bounds the variational error).
"""
import json
import os

import numpy as np
import pytest

from oracle import chain, models

pytestmark = pytest.mark.gpu

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "ed_energies.json")))


@pytest.fixture(scope="module")
def ctx():
    import networksolvers_amd as ns
    c = ns.Context(0)
    yield c
    c.close()


def run_device(ctx, W, mps0, precision="fp64", env_precision="fp64", **kw):
    import networksolvers_amd as ns
    H = ns.TTNO(ctx, W)
    psi = ns.TTN.from_arrays(ctx, mps0)
    prob = ns.EigsolveProblem(state=psi, operator=ns.ProjTTN(H, env_precision))
    from networksolvers_amd.sweep import eigsolve_problem
    E, psi = eigsolve_problem(prob, nsites=2, updater_kwargs=dict(precision=precision), **kw)
    return E, psi, prob


def test_c1_heisenberg_n20_parity(ctx):
    N = 20
    W = models.heisenberg_mpo(N)
    mps0 = chain.random_mps(N, 2, 4, seed=0)
    maxdim = [10, 20, 40, 64, 64]
    ref = chain.dmrg_chain(W, mps0, nsweeps=5, maxdim=maxdim, cutoff=1e-10)
    E, psi, prob = run_device(ctx, W, mps0, nsweeps=5, inserter_kwargs=dict(trunc=dict(cutoff=1e-10, maxdim=maxdim)))
    log = prob.region_log
    assert len(log) == len(ref.log) == 5 * 2 * (N - 1)
    for (region, e, kept, terr, route), r in zip(log, ref.log):
        assert tuple(region) == r.region
        assert kept == r.kept, (region, kept, r.kept)
        assert route == r.route == "eigen"
        assert abs(e - r.energy) <= 1e-10 * abs(r.energy)
        # eigen-route truncation errors are sums of eigenvalues of rho, each known to ~eps * tr(rho)
        assert abs(terr - r.truncerr) <= 1e-6 * r.truncerr + 1e-13
    assert abs(E - ref.energy) <= 1e-10 * abs(ref.energy)
    assert abs(E - GOLDEN["heisenberg_s12_chain"]["20"]) < 1e-8
    # the returned state: same bond dimensions, energy as an FP64 expectation value
    import networksolvers_amd as ns
    assert psi.linkdims() == [a.shape[2] for a in ref.mps[:-1]]
    e_exp = ns.expect_mpo(psi, ns.TTNO(ctx, W))
    assert abs(e_exp - chain.mpo_expectation(W, ref.mps)) < 1e-10


def test_s1_chain_reference_constant(ctx):
    """examples/dmrg.jl:41-43: the S=1 N=10 chain's energy -12.8945601."""
    W = models.heisenberg_mpo(10, d=3)
    mps0 = chain.random_mps(10, 3, 4, seed=2)
    E, psi, _ = run_device(ctx, W, mps0, nsweeps=5,
                           inserter_kwargs=dict(trunc=dict(cutoff=1e-12, maxdim=[10, 40, 80, 160])))
    assert abs(E - (-12.8945601)) < 1e-7


@pytest.mark.parametrize("jz", [0.0, 0.5, 1.5])
def test_xxz_vs_ed(ctx, jz):
    N = 10
    W = models.xxz_mpo(N, Jz=jz)
    mps0 = chain.random_mps(N, 2, 4, seed=1)
    E, _, _ = run_device(ctx, W, mps0, nsweeps=6, inserter_kwargs=dict(trunc=dict(cutoff=1e-13, maxdim=64)))
    assert abs(E - GOLDEN["xxz_s12_chain"][f"10_{jz}"]) < 1e-10


def test_fp32_form_energy_error(ctx):
    """FP32 H_eff contractions with FP64 Krylov/factorize: |dE/E| reported vs FP64."""
    N = 20
    W = models.heisenberg_mpo(N)
    mps0 = chain.random_mps(N, 2, 4, seed=0)
    kw = dict(nsweeps=6, inserter_kwargs=dict(trunc=dict(cutoff=1e-12, maxdim=[16, 32, 64])))
    E64, _, _ = run_device(ctx, W, mps0, **kw)
    E32, psi32, _ = run_device(ctx, W, mps0, precision="fp32", **kw)
    import networksolvers_amd as ns
    e32 = ns.expect_mpo(psi32, ns.TTNO(ctx, W))
    print(f"fp32 form: Ritz dE/E={abs(E32 - E64) / abs(E64):.2e}  FP64-expectation dE/E={abs(e32 - E64) / abs(E64):.2e}")
    assert abs(e32 - E64) / abs(E64) < 1e-9


def test_chi256_region_parity(ctx):
    """One region step at chi = 256 (n = 262144) against the oracle."""
    import networksolvers_amd as ns
    N, lo, chi = 24, 11, 256
    W = models.heisenberg_mpo(N)
    mps = chain.random_mps(N, 2, chi, seed=4)
    mps = chain.orthogonalize_chain(mps, lo, lo + 1)
    L = np.ones((1, 1, 1))
    for i in range(lo):
        L = chain.env_left(L, mps[i], W[i])
    R = np.ones((1, 1, 1))
    for i in range(N - 1, lo + 1, -1):
        R = chain.env_right(R, mps[i], W[i])
    th = chain.theta2(mps[lo], mps[lo + 1])
    from oracle import linalg
    shape = th.shape
    e0, x0, _ = linalg.lanczos_eigsolve(
        lambda t: chain.heff2_apply(L, W[lo], W[lo + 1], R, t.reshape(shape)).reshape(-1), th.reshape(-1))
    Uo, Co, io = linalg.factorize(x0.reshape(shape[0] * 2, -1), cutoff=1e-10, maxdim=chi)
    H = ns.HeffMap(ctx.upload(L), ctx.upload(W[lo]), ctx.upload(W[lo + 1]), ctx.upload(R))
    x = ctx.upload(th)
    res = ns.lanczos_eig(H, x)
    assert abs(res.eigval - e0) < 1e-11 * abs(e0)
    U, C, info = ns.factorize(x, 2, False, cutoff=1e-10, maxdim=chi)
    assert info.kept == io.kept
    assert abs(info.truncerr - io.truncerr) < 1e-6 * io.truncerr + 1e-14
    np.testing.assert_allclose(U.numpy().reshape(Uo.shape) @ C.numpy().reshape(Co.shape), Uo @ Co, atol=1e-9)
