"""GPU parity of the densitymatrix subspace expansion on QN chains
(subspace/densitymatrix.jl on block-sparse tensors) against the oracle.

Bars: one expansion step (either side of the region) equals the oracle's: the
same new link sectors, and the expanded site, next site and local tensor
element-wise to 1e-10 (random couplings keep every sector's spectrum
non-degenerate, so the eigenvectors are unique up to the shared sign rule);
QN DMRG with the expansion equals the oracle region by region (kept dimension
exact, Ritz value 1e-10 relative, every link's sectors); the reference's QN
example reaches -12.8945601.
"""
import numpy as np
import pytest

from oracle import chain, models
from oracle import qn as oqn

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import networksolvers_amd as ns
    c = ns.Context(0)
    yield c
    c.close()


def _neel_s1(N):
    return [0 if (i + 1) % 2 == 0 else 2 for i in range(N)]


def _disordered_xxz_s1(N, seed):
    """S=1 XXZ chain (Jz = 0.7 Jxy), bond couplings in [0.7, 1.3], fields in [-0.2, 0.2]."""
    rng = np.random.default_rng(seed)
    W = models.xxz_mpo(N, d=3, Jxy=1.0, Jz=0.7, hz=1.0)
    for i in range(N):
        r = 0 if i == 0 else 4
        W[i][r, 0] *= rng.uniform(-0.2, 0.2)
        if i < N - 1:
            W[i][r, 1:4] *= rng.uniform(0.7, 1.3)
    return W


def _gauge(mps, links, qs, lo, hi):
    """Sites left of lo left-orthonormal and right of hi right-orthonormal (QR on the
    dense tensors; QR keeps the charge blocks up to round-off, which is cleared)."""
    N = len(mps)
    mps = [m.copy() for m in mps]

    def clean(i):
        ql, qr = links[i].charges(), links[i + 1].charges()
        allowed = ql[:, None, None] + qs[None, :, None] == qr[None, None, :]
        mps[i][~allowed] = 0.0

    for i in range(lo):
        A = mps[i].reshape(-1, mps[i].shape[2])
        Q, R = np.linalg.qr(A)
        mps[i] = Q.reshape(mps[i].shape)
        mps[i + 1] = np.tensordot(R, mps[i + 1], axes=([1], [0]))
        clean(i)
    for i in range(N - 1, hi, -1):
        B = mps[i].reshape(mps[i].shape[0], -1)
        Q, R = np.linalg.qr(B.T)
        mps[i] = Q.T.reshape(mps[i].shape)
        mps[i - 1] = np.tensordot(mps[i - 1], R.T, axes=([2], [0]))
        clean(i)
    for i in range(N):
        clean(i)
    return mps


def _device_state(ctx, mps, links):
    from networksolvers_amd import qn
    site = qn.spin_site_index(3)
    ql = [qn.QNIndex(lk.sectors) for lk in links]
    tensors = [qn.QNTensor.from_dense(ctx, m, [ql[i], site, ql[i + 1]], [1, 1, -1], atol=1e-14)
               for i, m in enumerate(mps)]
    return qn.QNMPS(ctx, tensors, ql, site), site


@pytest.mark.parametrize("left", [True, False])
def test_qn_expansion_step_vs_oracle(ctx, left):
    import networksolvers_amd as ns
    from networksolvers_amd import qn
    N = 10
    W = _disordered_xxz_s1(N, 3)
    qs = oqn.spin_charges(3)
    res = oqn.qn_dmrg_chain(W, _neel_s1(N), 3, nsweeps=2, maxdim=[6, 12], cutoff=1e-10)
    links = list(res.links)
    lo, hi = (4, 5) if left else (3, 4)
    mps = _gauge(res.mps, links, qs, lo, hi)
    prev = [lo - 1, lo] if left else [hi, hi + 1]
    th = chain.theta2(mps[lo], mps[hi])
    # oracle step
    o_mps, o_links = [m.copy() for m in mps], list(links)
    cW = oqn.mpo_link_charges(W, qs)
    o_th, changed = oqn.densitymatrix_expand_qn(o_mps, o_links, qs, cW, W, chain.ChainEnvs(W), prev, (lo, hi), th,
                                                cutoff=1e-10, maxdim=40, mindim=1, expansion_factor=1.5)
    assert changed
    # device step from the same tensors, the operator at the previous position
    psi, site = _device_state(ctx, mps, links)
    H = qn.QNMPO(ctx, W, site)
    P = ns.sweep.projected_operator(H)
    P.position(psi, prev)
    psi.ortho_region = [lo, hi]
    theta = qn.extract(psi, [lo, hi])
    np.testing.assert_allclose(theta.to_dense(), th, atol=1e-13)
    prob = ns.EigsolveProblem(state=psi, operator=P)
    prob, d_th = qn.densitymatrix_expand(prob, theta, [lo, hi], trunc=dict(cutoff=1e-10, maxdim=40, mindim=1),
                                         expansion_factor=1.5)
    p = prev[0] if left else prev[1]
    ia = p + 1 if left else p
    assert list(zip(psi.links[ia].q, psi.links[ia].n)) == [tuple(x) for x in o_links[ia].sectors]
    assert psi.links[ia].dim > sum(n for _, n in links[ia].sectors)
    v = lo if left else hi
    np.testing.assert_allclose(psi[p].to_dense(), o_mps[p], atol=1e-10)
    np.testing.assert_allclose(psi[v].to_dense(), o_mps[v], atol=1e-10)
    np.testing.assert_allclose(d_th.to_dense(), o_th, atol=1e-10)


def test_qn_dmrg_densitymatrix_vs_oracle(ctx):
    """QN DMRG with the expansion on the disordered spin-1 chain, region by region."""
    import networksolvers_amd as ns
    from networksolvers_amd import qn
    N = 12
    W = _disordered_xxz_s1(N, 7)
    site = qn.spin_site_index(3)
    prob = ns.EigsolveProblem(state=qn.qn_product_mps(ctx, _neel_s1(N), site),
                              operator=ns.sweep.projected_operator(qn.QNMPO(ctx, W, site)))
    trunc = dict(cutoff=1e-8, maxdim=[8, 16, 32])
    E, psi = ns.sweep.eigsolve_problem(prob, nsweeps=4, nsites=2, inserter_kwargs=dict(trunc=trunc),
                                       extracter_kwargs=dict(trunc=trunc, subspace_algorithm="densitymatrix",
                                                             expansion_factor=1.2))
    ref = oqn.qn_dmrg_chain(W, _neel_s1(N), 3, nsweeps=4, maxdim=[8, 16, 32], cutoff=1e-8,
                            subspace_algorithm="densitymatrix", expansion_factor=1.2)
    assert len(prob.region_log) == len(ref.log)
    for a, b in zip(prob.region_log, ref.log):
        assert a[2] == b[2]
        assert abs(a[1] - b[1]) < 1e-10 * abs(b[1])
    for i in range(N + 1):
        assert list(zip(psi.links[i].q, psi.links[i].n)) == [tuple(x) for x in ref.links[i].sectors]
    assert abs(E - ref.energy) < 1e-10 * abs(E)


def test_qn_densitymatrix_reference_example(ctx):
    """examples/dmrg.jl:10-43 with conserve_qns=true on the device."""
    import networksolvers_amd as ns
    from networksolvers_amd import qn
    N = 10
    W = models.heisenberg_mpo(N, d=3)
    site = qn.spin_site_index(3)
    H = qn.QNMPO(ctx, W, site)
    prob = ns.EigsolveProblem(state=qn.qn_product_mps(ctx, _neel_s1(N), site),
                              operator=ns.sweep.projected_operator(H))
    trunc = dict(cutoff=1e-12, maxdim=[10, 40, 80, 160])
    E, psi = ns.sweep.eigsolve_problem(prob, nsweeps=5, nsites=2, inserter_kwargs=dict(trunc=trunc),
                                       extracter_kwargs=dict(trunc=trunc, subspace_algorithm="densitymatrix",
                                                             expansion_factor=1.1))
    ref = oqn.qn_dmrg_chain(W, _neel_s1(N), 3, nsweeps=5, maxdim=[10, 40, 80, 160], cutoff=1e-12,
                            subspace_algorithm="densitymatrix", expansion_factor=1.1)
    assert abs(E - (-12.8945601)) < 1e-7
    assert abs(E - ref.energy) < 1e-10 * abs(E)
    assert abs(ns.expect_mpo(psi, H) - E) < 1e-9
    assert max(psi.linkdims()) > 99        # plain 2-site DMRG stops at 99 on this run


@pytest.mark.parametrize("right", [True, False])
def test_qn_qr_step_vs_oracle(ctx, right):
    """The block-wise QR step of the walk equals the oracle's (same bond sectors,
    site and neighbour tensors element-wise: the thin QR with non-negative
    diag(R) is unique for full-rank sectors)."""
    from networksolvers_amd import qn
    N = 8
    W = _disordered_xxz_s1(N, 11)
    qs = oqn.spin_charges(3)
    res = oqn.qn_dmrg_chain(W, _neel_s1(N), 3, nsweeps=2, maxdim=[6, 12], cutoff=1e-10)
    x, y = (3, 4) if right else (4, 3)
    o_mps, o_links = [m.copy() for m in res.mps], list(res.links)
    oqn.qn_qr_toward(o_mps, o_links, qs, x, y)
    psi, site = _device_state(ctx, res.mps, res.links)
    psi.ortho_region = [x]
    qn.qr_toward(psi, x, y)
    bond = x + 1 if right else x
    assert list(zip(psi.links[bond].q, psi.links[bond].n)) == [tuple(t) for t in o_links[bond].sectors]
    np.testing.assert_allclose(psi[x].to_dense(), o_mps[x], atol=1e-12)
    np.testing.assert_allclose(psi[y].to_dense(), o_mps[y], atol=1e-12)
    assert psi.ortho_region == [y]


def test_qn_heff1_apply(ctx):
    """The 1-site QN H_eff equals the dense contraction of the same tensors."""
    from networksolvers_amd import qn
    N = 8
    W = _disordered_xxz_s1(N, 13)
    res = oqn.qn_dmrg_chain(W, _neel_s1(N), 3, nsweeps=2, maxdim=[6, 12], cutoff=1e-10)
    psi, site = _device_state(ctx, res.mps, res.links)
    H = qn.QNMPO(ctx, W, site)
    P = qn.QNProjTTN(H, ctx)
    v = 4
    P.position(psi, [v])
    op = qn.heff_map(P, psi)
    x = psi[v].copy()
    y = psi[v].copy()
    op.run(x.data, y.data)
    envs = chain.ChainEnvs(W)
    ref = chain.heff1_apply(envs.get(res.mps, v - 1, v), W[v], envs.get(res.mps, v + 1, v), res.mps[v])
    assert np.linalg.norm(y.to_dense() - ref) < 1e-12 * np.linalg.norm(ref)
    op.close()


def test_qn_1site_densitymatrix_vs_oracle(ctx):
    """1-site QN DMRG with the densitymatrix expansion (max_expand 4, the README's
    1-site form): region by region the device equals the oracle (Ritz value 1e-10
    relative), and so do the final link sectors."""
    import networksolvers_amd as ns
    from networksolvers_amd import qn
    N = 10
    W = _disordered_xxz_s1(N, 17)
    site = qn.spin_site_index(3)
    prob = ns.EigsolveProblem(state=qn.qn_product_mps(ctx, _neel_s1(N), site),
                              operator=ns.sweep.projected_operator(qn.QNMPO(ctx, W, site)))
    trunc = dict(cutoff=1e-8, maxdim=[8, 16, 32])
    energies = []
    E, psi = ns.sweep.eigsolve_problem(prob, nsweeps=4, nsites=1, inserter_kwargs=dict(trunc=trunc),
                                       extracter_kwargs=dict(trunc=trunc, subspace_algorithm="densitymatrix",
                                                             max_expand=4),
                                       region_callback=lambda problem, **kw: energies.append(problem.eigenvalue))
    ref = oqn.qn_dmrg_chain(W, _neel_s1(N), 3, nsites=1, nsweeps=4, maxdim=[8, 16, 32], cutoff=1e-8,
                            subspace_algorithm="densitymatrix", max_expand=4)
    assert len(energies) == len(ref.log)
    for e, r_ in zip(energies, ref.log):
        assert abs(e - r_[1]) < 1e-10 * abs(r_[1])
    for i in range(N + 1):
        assert list(zip(psi.links[i].q, psi.links[i].n)) == [tuple(t) for t in ref.links[i].sectors]
    assert max(psi.linkdims()) > 1
