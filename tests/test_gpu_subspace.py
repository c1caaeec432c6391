"""GPU parity of the densitymatrix subspace expansion (subspace/densitymatrix.jl)
inside 1- and 2-site DMRG, on trees and chains, against the oracle.

Bars: final energy to 1e-10 relative of the oracle's, identical bond dimensions,
and the reference's own targets (test_tree_dmrg.jl:78-90: |E - E_ED| < 1e-5;
examples/dmrg.jl:41-43: -12.8945601).
"""
import json
import os

import numpy as np
import pytest

from oracle import chain as ochain
from oracle import models as omodels
from oracle import tree as otree

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import networksolvers_amd as ns
    c = ns.Context(0)
    yield c
    c.close()


def _ed(model, key):
    with open(os.path.join(os.path.dirname(__file__), "golden", "ed_energies.json")) as f:
        return json.load(f)[model][key]


def build_tree():
    V = [(0, 0)] + [(b, s) for b in range(1, 4) for s in range(1, 4)]
    E = []
    for b in range(1, 4):
        E.append(((0, 0), (b, 1)))
        for s in range(2, 4):
            E.append(((b, s - 1), (b, s)))
    return V, E


def neel(T, V, d=2):
    psi = {}
    for j, v in enumerate(V, start=1):
        k = len(T.adj[v])
        a = np.zeros([1] * k + [d])
        a[(0,) * k + (0 if j % 2 == 0 else d - 1,)] = 1.0
        psi[v] = a
    return psi


def _path_layout(Wc, mps):
    """Chain MPO/MPS (W[wl, wr, s', s], A[l, s, r]) in the oracle's path-graph layout."""
    N = len(Wc)
    W = {0: Wc[0][0], N - 1: Wc[N - 1][:, 0]}
    W.update({i: Wc[i] for i in range(1, N - 1)})
    psi = {0: mps[0][0].T.copy(), N - 1: mps[N - 1][:, :, 0].copy()}
    psi.update({i: np.transpose(mps[i], (0, 2, 1)).copy() for i in range(1, N - 1)})
    return W, psi


def test_tree_1site_densitymatrix_reference_case(ctx):
    import networksolvers_amd as ns
    from networksolvers_amd.sweep import EigsolveProblem, eigsolve_problem, projected_operator
    V, E = build_tree()
    T = otree.Tree(V, E)
    G = ns.Graph(V, E)
    W = ns.heisenberg_tree_mpo(G)
    psi0 = neel(T, V)
    trunc = dict(cutoff=1e-5, maxdim=40)
    prob = EigsolveProblem(state=ns.TreeTTN.from_arrays(ctx, G, psi0), operator=projected_operator(ns.TreeTTNO(ctx, G, W)))
    En, psi = eigsolve_problem(prob, nsweeps=5, nsites=1,
                               extracter_kwargs=dict(trunc=trunc, subspace_algorithm="densitymatrix"),
                               inserter_kwargs=dict(trunc=trunc))
    ex = _ed("tree", "build_tree_3_3_heisenberg_sz0")
    assert abs(En - ex) < 1e-5
    ref = otree.tree_dmrg(V, E, W, psi0, nsweeps=5, nsites=1, maxdim=40, cutoff=1e-5,
                          subspace_algorithm="densitymatrix")
    assert abs(En - ref.energy) < 1e-10 * abs(ex)
    for a, b in E:
        assert psi.linkdim(a, b) == ref.psi[a].shape[T.adj[a].index(b)]
    for (rg, e, *_), (rr, er, *_) in zip(prob.region_log, ref.log):
        assert tuple(rg) == tuple(rr)


def test_chain_2site_densitymatrix_reference_example(ctx):
    """examples/dmrg.jl:10-43: S=1 Heisenberg N=10, 2-site DMRG with densitymatrix
    expansion (factor 1.1), cutoff 1e-12, maxdim [10,40,80,160], 5 sweeps."""
    import networksolvers_amd as ns
    N = 10
    Wc = omodels.heisenberg_mpo(N, d=3)
    mps0 = ochain.product_mps([0 if (i + 1) % 2 == 0 else 2 for i in range(N)], 3)
    trunc = dict(cutoff=1e-12, maxdim=[10, 40, 80, 160])
    En, psi = ns.dmrg(ns.TTNO(ctx, Wc), ns.TTN.from_arrays(ctx, mps0), nsweeps=5, nsites=2,
                      extracter_kwargs=dict(trunc=trunc, subspace_algorithm="densitymatrix", expansion_factor=1.1),
                      inserter_kwargs=dict(trunc=trunc))
    assert abs(En - (-12.8945601)) < 1e-7
    W, p0 = _path_layout(Wc, mps0)
    ref = otree.tree_dmrg(list(range(N)), [(i, i + 1) for i in range(N - 1)], W, p0, nsweeps=5, nsites=2,
                          maxdim=[10, 40, 80, 160], cutoff=1e-12, subspace_algorithm="densitymatrix",
                          expansion_factor=1.1)
    assert abs(En - ref.energy) < 1e-10 * abs(En)
    assert psi.linkdims() == [ref.psi[i].shape[-2 if 0 < i else 0] for i in range(N - 1)]


def test_chain_1site_densitymatrix_from_product_state(ctx):
    """1-site DMRG can only grow bonds through the expansion: from a Neel product
    state on a Heisenberg N=12 chain it reaches the ED energy and follows the
    oracle (energies 1e-10, bond dimensions exact)."""
    import networksolvers_amd as ns
    N = 12
    Wc = omodels.heisenberg_mpo(N)
    mps0 = ochain.product_mps([0, 1] * (N // 2), 2)
    trunc = dict(cutoff=1e-12, maxdim=64)
    En, psi = ns.dmrg(ns.TTNO(ctx, Wc), ns.TTN.from_arrays(ctx, mps0), nsweeps=8, nsites=1,
                      extracter_kwargs=dict(trunc=trunc, subspace_algorithm="densitymatrix"),
                      inserter_kwargs=dict(trunc=trunc))
    assert abs(En - _ed("heisenberg_s12_chain", "12")) < 1e-8
    W, p0 = _path_layout(Wc, mps0)
    ref = otree.tree_dmrg(list(range(N)), [(i, i + 1) for i in range(N - 1)], W, p0, nsweeps=8, nsites=1,
                          maxdim=64, cutoff=1e-12, subspace_algorithm="densitymatrix")
    assert abs(En - ref.energy) < 1e-10 * abs(En)
    assert psi.linkdims() == [ref.psi[i].shape[-2 if 0 < i else 0] for i in range(N - 1)]


def test_unknown_subspace_algorithm_raises(ctx):
    import networksolvers_amd as ns
    N = 4
    Wc = omodels.heisenberg_mpo(N)
    mps0 = ochain.product_mps([0, 1] * 2, 2)
    with pytest.raises(NotImplementedError, match="Subspace expansion"):
        ns.dmrg(ns.TTNO(ctx, Wc), ns.TTN.from_arrays(ctx, mps0), nsweeps=1, nsites=2,
                extracter_kwargs=dict(subspace_algorithm="ortho"))
