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

from oracle import chain, models
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
