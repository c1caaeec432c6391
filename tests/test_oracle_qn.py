"""Pins of the U(1) quantum-number restatement (CPU only).

* factorize_sectors without truncation reproduces every sector matrix, and its
  merged spectrum equals the spectrum of the block-diagonal matrix computed densely;
* the block-sparse cut: entries above docut are kept per sector, and a
  near-degenerate pair that straddles the cut is dropped together;
* QN DMRG of the spin-1 chain N=10 reaches the reference's constant -12.8945601
  (examples/dmrg.jl:41-43) and the ED fixture; for the S=1/2 chain N=12 the QN
  and the dense DMRG give the same energy and bond dimensions.
"""
import json
import os

import numpy as np
import pytest
import scipy.linalg as sl

from oracle import chain, linalg, models
from oracle import qn as oqn


def _ed(model, key):
    with open(os.path.join(os.path.dirname(__file__), "golden", "ed_energies.json")) as f:
        return json.load(f)[model][key]


@pytest.mark.parametrize("rows_trailing", [False, True])
@pytest.mark.parametrize("route", ["svd", "eigen"])
def test_sectors_without_truncation(rows_trailing, route):
    rng = np.random.default_rng(1)
    mats = [rng.standard_normal((5, 7)), rng.standard_normal((3, 2)), rng.standard_normal((8, 8))]
    res, kept, te = oqn.factorize_sectors(mats, rows_trailing, cutoff=0.0, which_decomp=route)
    assert te < 1e-15            # the eigen route may drop a round-off eigenvalue of a rank-deficient Gram matrix
    for M, (U, C) in zip(mats, res):
        np.testing.assert_allclose(U @ C if not rows_trailing else C @ U, M, atol=1e-12)
    spec = np.sort(np.concatenate([np.linalg.svd(M, compute_uv=False) ** 2 for M in mats]))[::-1]
    dense = np.linalg.svd(sl.block_diag(*mats), compute_uv=False) ** 2
    dense = np.sort(dense)[::-1]
    np.testing.assert_allclose(spec, dense[:len(spec)], rtol=1e-12, atol=1e-14)   # the rest are zero
    assert np.all(dense[len(spec):] < 1e-24)
    assert sum(kept) == sum(min(M.shape) for M in mats)


def _diag_sector(vals, rng):
    n = len(vals)
    Q1, _ = np.linalg.qr(rng.standard_normal((n, n)))
    Q2, _ = np.linalg.qr(rng.standard_normal((n, n)))
    return Q1 @ np.diag(np.sqrt(vals)) @ Q2.T


def test_cut_keeps_entries_above_docut_per_sector():
    rng = np.random.default_rng(2)
    a = _diag_sector([1.0, 0.3, 1e-3], rng)
    b = _diag_sector([0.5, 2e-3, 1e-6], rng)
    _, kept, _ = oqn.factorize_sectors([a, b], False, maxdim=4, which_decomp="svd")
    assert kept == [2, 2]                     # 1, 0.5, 0.3, 2e-3 kept; 1e-3 and 1e-6 dropped
    # near-degenerate pair across sectors at the cut: both dropped (docut nudged above them)
    c = _diag_sector([1.0, 1e-2], rng)
    d = _diag_sector([0.5, 1e-2 * (1 + 1e-5)], rng)
    _, kept, _ = oqn.factorize_sectors([c, d], False, maxdim=3, which_decomp="svd")
    assert kept == [1, 1]


def test_qn_dmrg_spin1_reference_constant():
    N = 10
    W = models.heisenberg_mpo(N, d=3)
    res = oqn.qn_dmrg_chain(W, [0 if (i + 1) % 2 == 0 else 2 for i in range(N)], 3, nsweeps=6,
                            maxdim=[10, 40, 80, 160, 243], cutoff=1e-12)
    assert abs(res.energy - (-12.8945601)) < 1e-7
    assert abs(res.energy - _ed("heisenberg_s1_chain", "10")) < 1e-9
    assert all(q == 0 for q, _ in res.links[-1].sectors)          # total Sz = 0
    assert all(sum(n for _, n in res.links[i + 1].sectors) == res.mps[i].shape[2] for i in range(N - 1))


def test_qn_equals_dense_dmrg():
    N = 12
    W = models.heisenberg_mpo(N)
    states = [0, 1] * (N // 2)
    q = oqn.qn_dmrg_chain(W, states, 2, nsweeps=6, maxdim=[8, 16, 32, 64], cutoff=1e-10)
    d = chain.dmrg_chain(W, chain.product_mps(states, 2), nsweeps=6, maxdim=[8, 16, 32, 64], cutoff=1e-10)
    assert abs(q.energy - d.energy) < 1e-12
    assert [m.shape[2] for m in q.mps[:-1]] == [m.shape[2] for m in d.mps[:-1]]


def _neel_s1(N):
    """The reference example's Neel start: "Up" (m = +1) on even 1-based sites, "Dn" (m = -1) otherwise."""
    return [0 if (i + 1) % 2 == 0 else 2 for i in range(N)]


def _path_dmrg_dense(W, states, d, **kw):
    from oracle import tree as otree
    N = len(W)
    V, E = list(range(N)), [(i, i + 1) for i in range(N - 1)]
    Wd = {0: W[0][0], N - 1: W[N - 1][:, 0]}
    Wd.update({i: W[i] for i in range(1, N - 1)})
    psi = {}
    for i, s in enumerate(states):
        k = 1 if i in (0, N - 1) else 2
        a = np.zeros([1] * k + [d])
        a[(0,) * k + (s,)] = 1.0
        psi[i] = a
    return otree.tree_dmrg(V, E, Wd, psi, nsites=2, **kw)


def test_qn_densitymatrix_reference_example():
    """examples/dmrg.jl:10-43 with conserve_qns=true: S=1 chain N=10 from the Neel
    state, 2-site DMRG with densitymatrix expansion (factor 1.1, cutoff 1e-12,
    maxdim [10,40,80,160], 5 sweeps) reaches the reference's -12.8945601 and the ED
    fixture, and the expansion grows the links beyond plain 2-site DMRG's."""
    N = 10
    W = models.heisenberg_mpo(N, d=3)
    q = oqn.qn_dmrg_chain(W, _neel_s1(N), 3, nsweeps=5, maxdim=[10, 40, 80, 160], cutoff=1e-12,
                          subspace_algorithm="densitymatrix", expansion_factor=1.1)
    assert abs(q.energy - (-12.8945601)) < 1e-7
    assert abs(q.energy - _ed("heisenberg_s1_chain", "10")) < 1e-9
    plain = oqn.qn_dmrg_chain(W, _neel_s1(N), 3, nsweeps=5, maxdim=[10, 40, 80, 160], cutoff=1e-12)
    grown = [sum(n for _, n in lk.sectors) for lk in q.links]
    assert max(grown) > max(sum(n for _, n in lk.sectors) for lk in plain.links)


def _disordered_xxz_s1(N, seed):
    """S=1 XXZ chain (Jz = 0.7 Jxy) with bond couplings J_i in [0.7, 1.3] and fields
    h_i in [-0.2, 0.2]: the uniform MPO with its coupling row scaled per site."""
    rng = np.random.default_rng(seed)
    W = models.xxz_mpo(N, d=3, Jxy=1.0, Jz=0.7, hz=1.0)
    for i in range(N):
        r = 0 if i == 0 else 4
        W[i][r, 0] *= rng.uniform(-0.2, 0.2)
        if i < N - 1:
            W[i][r, 1:4] *= rng.uniform(0.7, 1.3)
    return W


@pytest.mark.parametrize("factor,maxdim", [(1.1, [10, 20, 40]), (1.5, [8, 16, 24])])
def test_qn_expansion_equals_dense_expansion(factor, maxdim):
    """The expansion's S is block diagonal in the charge of the expanded link, so
    the sector-wise eigen-decomposition with cross-sector truncation keeps the same
    space as the dense one: region by region the QN sweep equals the dense
    expansion on the path graph (kept dimension exact, Ritz value 1e-10).  The
    chain is an S=1 XXZ chain with random bond couplings and fields, so that no cut
    meets a (near-)degenerate pair: there dense truncation keeps part of the pair
    and the block-sparse cut keeps or drops it whole (NDTensors' docut).  The
    cutoff (1e-8) sits well above the eigensolvers' absolute precision."""
    N = 10
    W = _disordered_xxz_s1(N, 5)
    kw = dict(nsweeps=4, maxdim=maxdim, cutoff=1e-8, subspace_algorithm="densitymatrix", expansion_factor=factor)
    q = oqn.qn_dmrg_chain(W, _neel_s1(N), 3, **kw)
    d = _path_dmrg_dense(W, _neel_s1(N), 3, **kw)
    assert len(q.log) == len(d.log)
    for a, b in zip(q.log, d.log):
        assert a[2] == b[2]
        assert abs(a[1] - b[1]) < 1e-10 * abs(b[1])


@pytest.mark.parametrize("nsites", [2, 1])
@pytest.mark.parametrize("left", [True, False])
def test_qn_expansion_step_invariants(left, nsites):
    """One expansion step on a converged-ish state (2-site region, or 1-site region
    with the local tensor the site itself): [A | U] is an isometry, the product
    A.theta is unchanged (expander = [A|U]^T A keeps the state), each link charge
    gains at most its own new vectors, and every tensor keeps zero weight outside
    the charge-allowed blocks."""
    N = 8
    W = models.heisenberg_mpo(N, d=3)
    qs = oqn.spin_charges(3)
    res = oqn.qn_dmrg_chain(W, _neel_s1(N), 3, nsweeps=2, maxdim=6, cutoff=1e-12)
    mps, links = [m.copy() for m in res.mps], list(res.links)
    cW = oqn.mpo_link_charges(W, qs)
    envs = chain.ChainEnvs(W)
    lo, hi = (4, 5) if left else (2, 3)
    if nsites == 1:
        lo, hi = (lo, lo) if left else (hi, hi)
    p = lo - 1 if left else hi + 1
    ia = p + 1 if left else p
    # put the gauge centre on the region, as the extracter does: p is an isometry toward it
    for i in range(N - 1):
        if i < lo:
            A = mps[i].reshape(-1, mps[i].shape[2])
            Q, Rm = np.linalg.qr(A)
            if Q.shape[1] == A.shape[1]:
                mps[i] = Q.reshape(mps[i].shape)
                mps[i + 1] = np.tensordot(Rm, mps[i + 1], axes=([1], [0]))
    for i in range(N - 1, hi, -1):
        B = mps[i].reshape(mps[i].shape[0], -1)
        Q, Rm = np.linalg.qr(B.T)
        if Q.shape[1] == B.shape[0]:
            mps[i] = Q.T.reshape(mps[i].shape)
            mps[i - 1] = np.tensordot(mps[i - 1], Rm.T, axes=([2], [0]))
    th = chain.theta2(mps[lo], mps[hi]) if nsites == 2 else mps[lo].copy()
    A0, old = mps[p].copy(), links[ia]
    last = th.ndim - 1
    before = (np.tensordot(A0, th, axes=([2], [0])) if left else np.tensordot(th, A0, axes=([last], [0])))
    prev = ([lo - 1, lo] if left else [hi, hi + 1]) if nsites == 2 else [p]
    th2, changed = oqn.densitymatrix_expand_qn(mps, links, qs, cW, W, envs, prev, (lo, hi), th, cutoff=1e-12,
                                               maxdim=64, mindim=1)
    assert sorted(changed) == sorted([p, lo if left else hi])
    A1 = mps[p]
    Am = A1.reshape(-1, A1.shape[2]) if left else A1.reshape(A1.shape[0], -1).T
    np.testing.assert_allclose(Am.T @ Am, np.eye(Am.shape[1]), atol=1e-10)
    after = (np.tensordot(A1, th2, axes=([2], [0])) if left else np.tensordot(th2, A1, axes=([last], [0])))
    np.testing.assert_allclose(after, before, atol=1e-10)
    new = dict(links[ia].sectors)
    assert sum(new.values()) > sum(n for _, n in old.sectors)
    for q, n in old.sectors:
        assert new[q] >= n
    # charge-allowed support of the new site tensor (the dense QR gauge above leaves
    # round-off in forbidden blocks)
    ql, qa = (links[p].charges(), links[p + 1].charges())
    allowed = ql[:, None, None] + qs[None, :, None] == qa[None, None, :]
    assert np.max(np.abs(A1[~allowed]), initial=0.0) < 1e-14


@pytest.mark.parametrize("right", [True, False])
def test_qn_qr_step(right):
    """The block-wise QR of the orthogonalize walk (ITensors qr of a QN tensor):
    the two-site product is unchanged, the new site tensor is an isometry toward
    the moved centre, each bond charge keeps min(rows, cols) vectors, and with
    every sector tall (full column rank) Q and R equal the dense qr_pos ones."""
    rng = np.random.default_rng(3)
    d = 3
    qs = oqn.spin_charges(d)
    ll = oqn.Link([(-2, 2), (0, 3), (2, 2)])
    lm = oqn.Link([(-4, 1), (-2, 2), (0, 3), (2, 2), (4, 1)])
    lr = oqn.Link([(-2, 3), (0, 4), (2, 3)])

    def rand_site(a, b):
        qa, qb = a.charges(), b.charges()
        allowed = qa[:, None, None] + qs[None, :, None] == qb[None, None, :]
        return np.where(allowed, rng.standard_normal(allowed.shape), 0.0)

    mps = [rand_site(ll, lm), rand_site(lm, lr)]
    links = [ll, lm, lr]
    before = np.tensordot(mps[0], mps[1], axes=([2], [0]))
    x, y = (0, 1) if right else (1, 0)
    A0 = mps[x].copy()
    oqn.qn_qr_toward(mps, links, qs, x, y)
    after = np.tensordot(mps[0], mps[1], axes=([2], [0]))
    np.testing.assert_allclose(after, before, atol=1e-12)
    A = mps[x]
    M = A.reshape(-1, A.shape[2]) if right else A.reshape(A.shape[0], -1).T
    np.testing.assert_allclose(M.T @ M, np.eye(M.shape[1]), atol=1e-12)
    # rows per bond charge vs the old bond's sector sizes
    if right:
        qrow = (ll.charges()[:, None] + qs[None, :]).reshape(-1)
    else:
        qrow = (-qs[:, None] + lr.charges()[None, :]).reshape(-1)
    expect = [(q, min(int(np.sum(qrow == q)), n)) for q, n in lm.sectors if np.sum(qrow == q) > 0]
    assert links[1].sectors == expect
    if all(int(np.sum(qrow == q)) >= n for q, n in lm.sectors):
        M0 = A0.reshape(-1, A0.shape[2]) if right else A0.reshape(A0.shape[0], -1).T
        Qd, _ = linalg.qr_pos(M0)
        np.testing.assert_allclose(M, Qd, atol=1e-12)


def test_qn_1site_densitymatrix_reaches_ground_state():
    """1-site QN DMRG with the densitymatrix expansion (the reference README's
    1-site form, max_expand 4) from the Neel state: the expansion grows the links
    from dimension 1, the energy reaches the spin-1 N=10 ED value within the
    reference's bar for 1-site DMRG with expansion (1e-5, test_tree_dmrg.jl:67),
    and the last Ritz value equals <psi|H|psi>.  (The block-wise QR keeps at most
    min(rows, cols) vectors per charge, so the QN sweep has fewer directions to
    rotate into than the dense one, which keeps rank-deficient columns; here it
    stops at 1.2e-6 above the ED energy where the dense sweep reaches it.)"""
    N = 10
    W = models.heisenberg_mpo(N, d=3)
    q = oqn.qn_dmrg_chain(W, _neel_s1(N), 3, nsites=1, nsweeps=12, maxdim=[10, 40, 80, 160], cutoff=1e-12,
                          subspace_algorithm="densitymatrix", max_expand=4)
    assert abs(q.energy - _ed("heisenberg_s1_chain", "10")) < 1e-5
    assert abs(chain.mpo_expectation(W, q.mps) - q.energy) < 1e-10
    assert max(sum(n for _, n in lk.sectors) for lk in q.links) > 1
    assert all(q_ == 0 for q_, _ in q.links[-1].sectors)
