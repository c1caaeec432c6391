"""U(1) quantum-number DMRG on a chain (test infrastructure only).

Restates the reference's 2-site sweep on QN-conserving tensors (README.md QN
timings; ITensors QN ITensors under eigsolve.jl / inserter.jl) with DENSE numpy
tensors whose charge-forbidden blocks are zero: environments, the effective
Hamiltonian and the Lanczos solver are the dense chain oracle's (the zero blocks
stay exactly zero), and only the truncated split is block-wise -- NDTensors'
block-sparse svd/eigen: each charge sector of theta (rows (l, s1) and columns
(s2, r) with q_l + q_s1 = q_r - q_s2 = q) is decomposed on its own, the sector
spectra are merged and truncated together by truncate!, and each sector keeps
its entries above the cut value docut (so a near-degenerate pair at the cut is
kept or dropped together).  The new link lists its sectors in increasing charge,
each with its kept vectors in decreasing spectral order.

The densitymatrix subspace expansion (subspace/densitymatrix.jl) is restated
the same way: its S is block diagonal in the charge of the expanded link, so rho
is decomposed per sector and truncated across sectors.

Pinned by (tests/test_oracle_qn.py): without truncation the sector split
reproduces theta exactly and its merged spectrum equals the dense one; the QN
ground energy of the spin-1 chain N=10 equals the reference's constant
-12.8945601 and the ED fixture; QN and dense DMRG energies agree.  With the
expansion: the reference's QN example (examples/dmrg.jl, conserve_qns=true)
reaches -12.8945601 and the ED fixture; region by region the QN sweep equals the
dense expansion of the tree restatement on the path graph (kept dimension exact,
Ritz value 1e-10, on a disordered chain with no degenerate cut); one expansion
step keeps the two-site state, leaves [A | U] an isometry and the tensors
charge-block-sparse.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import linalg
from .chain import ChainEnvs, heff2_apply, theta2
from .plans import euler_sweep_chain, get_or_last


def spin_charges(d):
    """2m for m = S, S-1, ..., -S."""
    return np.array([int(round(2 * ((d - 1) / 2 - i))) for i in range(d)])


def mpo_link_charges(W, qs):
    """MPO link charges from the left boundary: c_wr = c_wl + q_s' - q_s."""
    c = [np.zeros(W[0].shape[0], dtype=int)]
    for w in W:
        cr = {}
        for a, b, sp, s in zip(*np.nonzero(w)):
            v = c[-1][a] + qs[sp] - qs[s]
            if cr.setdefault(b, v) != v:
                raise ValueError("MPO does not conserve the charge")
        c.append(np.array([cr.get(b, 0) for b in range(w.shape[1])]))
    return c


@dataclass
class Link:
    """Sectors (charge, dim) in increasing charge; dense basis = sectors in order."""
    sectors: list

    def charges(self):
        return np.concatenate([np.full(n, q) for q, n in self.sectors]) if self.sectors else np.zeros(0, int)

    def offsets(self):
        out, o = {}, 0
        for q, n in self.sectors:
            out[q] = o
            o += n
        return out


def _sector_decomp(M, route):
    """(P sorted as NDTensors sorts it, row vectors V (len(P) x m) of the isometry
    over M's rows, in that order)."""
    m, n = M.shape
    if route == "svd":
        u, s, vh = np.linalg.svd(M, full_matrices=False)
        return s.astype(float) ** 2, u.T
    w, v = np.linalg.eigh(M @ M.T)
    order = np.argsort(-np.abs(w), kind="stable")
    return w[order], v[:, order].T


def factorize_sectors(mats, rows_trailing, *, cutoff=0.0, maxdim=None, mindim=1, which_decomp=None):
    """Block-diagonal truncated split of the sector matrices ``mats`` (list of
    2-D arrays).  Returns (list of (U_q, C_q) or None per sector, kept list,
    truncerr): U_q is rows x k (k x cols when rows_trailing) with the largest-
    magnitude entry of each isometry vector positive; C_q = U_q^T M (M U_q^T)."""
    route = linalg.choose_route(cutoff, which_decomp)
    dec = []
    for M in mats:
        Mi = M.T if rows_trailing else M
        dec.append(_sector_decomp(Mi, route))
    full = sum(min(M.shape) for M in mats)
    mx = full if maxdim is None else min(int(maxdim), full)
    key = (lambda v: abs(v)) if route == "eigen" else (lambda v: v)
    ent = [(key(p), p, q, i) for q, (P, _) in enumerate(dec) for i, p in enumerate(P)]
    ent.sort(key=lambda e: -e[0])           # stable: sector order, then spectral order, on ties
    P = np.array([e[1] for e in ent], dtype=float)
    n, truncerr, docut = linalg.truncate(P, mindim=mindim, maxdim=mx, cutoff=cutoff)
    Pp = P.copy()                           # the values truncate! compares (sign flip, negative tail zeroed)
    if len(Pp) and Pp[0] < 0:
        Pp = -Pp
    for j in range(len(Pp) - 1, -1, -1):
        if Pp[j] >= 0:
            break
        Pp[j] = 0.0
    kept = [0] * len(mats)
    if n < len(P):
        for j, e in enumerate(ent):
            if Pp[j] > docut:
                kept[e[2]] += 1
    else:
        kept = [len(d[0]) for d in dec]
    if sum(kept) == 0 and ent:
        kept[ent[0][2]] = 1
    out = []
    for q, M in enumerate(mats):
        k = kept[q]
        if k == 0:
            out.append(None)
            continue
        V = dec[q][1][:k].T                 # iso x k
        V, _ = linalg.fix_column_phases(V)
        if not rows_trailing:
            out.append((V, V.T @ M))
        else:
            out.append((V.T, M @ V))
    return out, kept, truncerr


def qn_split(theta, ql, qs, qr_link, rows_trailing, **trunc):
    """The split of a dense theta[l, s1, s2, r] by charge sectors: returns the
    dense left tensor [l, s1, m], right tensor [m, s2, r], the new link, truncerr."""
    cl, d, _, cr = theta.shape
    lo = ql.offsets()
    ro = qr_link.offsets()
    rows, cols = {}, {}
    for (qa, na) in ql.sectors:
        for s1 in range(d):
            rows.setdefault(qa + qs[s1], []).append((qa, s1, na))
    for (qb, nb) in qr_link.sectors:
        for s2 in range(d):
            cols.setdefault(qb - qs[s2], []).append((s2, qb, nb))
    sectors = sorted(q for q in rows if q in cols)
    mats, ridx, cidx = [], [], []
    for q in sectors:
        # rows sorted by (sector of l, s1), then the index inside the sector; columns by (s2, sector of r)
        ri = []
        for (qa, s1, na) in sorted(rows[q], key=lambda t: (ql.sectors.index((t[0], t[2])), t[1])):
            ri += [(lo[qa] + i, s1) for i in range(na)]
        ci = []
        for (s2, qb, nb) in sorted(cols[q], key=lambda t: (t[0], qr_link.sectors.index((t[1], t[2])))):
            ci += [(s2, ro[qb] + j) for j in range(nb)]
        ra, rs = np.array([x[0] for x in ri]), np.array([x[1] for x in ri])
        cs, cb = np.array([x[0] for x in ci]), np.array([x[1] for x in ci])
        M = theta[ra[:, None], rs[:, None], cs[None, :], cb[None, :]]
        mats.append(M)
        ridx.append(ri)
        cidx.append(ci)
    res, kept, truncerr = factorize_sectors(mats, rows_trailing, **trunc)
    mid = Link([(int(q), int(k)) for q, k in zip(sectors, kept) if k > 0])
    K = sum(k for _, k in mid.sectors)
    left = np.zeros((cl, d, K))
    right = np.zeros((K, d, cr))
    mo = mid.offsets()
    for q, r_, ri, ci in zip(sectors, res, ridx, cidx):
        if r_ is None:
            continue
        U, Cq = r_
        Lq, Rq = (Cq, U) if rows_trailing else (U, Cq)      # left: rows x k, right: k x cols
        k = Lq.shape[1]
        ra, rs = np.array([x[0] for x in ri]), np.array([x[1] for x in ri])
        cs, cb = np.array([x[0] for x in ci]), np.array([x[1] for x in ci])
        left[ra, rs, mo[q]:mo[q] + k] = Lq
        right[mo[q]:mo[q] + k, cs, cb] = Rq
    return left, right, mid, truncerr


def densitymatrix_expand_qn(mps, links, qs, cW, W, envs, prev_pos, region, local, *, cutoff, maxdim, mindim,
                            expansion_factor=1.5, max_expand=None, north_pass=1):
    """subspace/densitymatrix.jl:5-74 on a QN chain (the 2-site region's previous
    position p next to it): S = (1 - P_A)(E_p . W_p . A_p) is block diagonal in the
    charge Q of the link between p and the region, so rho = S S^T is decomposed per
    sector and truncated across sectors (the block-sparse eigen of a QN ITensor,
    eigen route, maxdim = compute_expansion); U is projected off A once more and
    checked (|U^T A| <= 1e-10, else nothing changes); the link gains, in every
    charge, the old basis followed by that charge's new vectors, A <- [A | U], and
    the next site and the local tensor are multiplied by expander = [A | U]^T A.
    Dense tensors throughout, charge-forbidden blocks zero.  Returns (local,
    changed sites)."""
    from .tree import compute_expansion
    lo, hi = min(region), max(region)
    prev = [v for v in prev_pos if v not in (lo, hi)]
    if len(prev) != 1 or prev[0] not in (lo - 1, hi + 1):
        return local, []
    p = prev[0]
    left = p == lo - 1                      # A = mps[p] [l, s, a]; otherwise [a, s, r]
    v = lo if left else hi
    A = mps[p]
    ia = p + 1 if left else p               # links[ia] is the link between p and v
    if left:
        E = envs.get(mps, p - 1, p)                                       # [bl, wl, l]
        S = np.einsum("xwl,lsa,wvts->xtav", E, A, W[p])                 # [bl, t, a, wa]
        Ar = A                                                            # rows (l, s)
        qrow = links[p].charges()[:, None] + qs[None, :]
        qcol = links[ia].charges()[:, None] + cW[p + 1][None, :]
    else:
        E = envs.get(mps, p + 1, p)                                       # [br, vr, r]
        S = np.einsum("asr,yvr,wvts->tyaw", A, E, W[p])                 # [t, br, a, wa]
        Ar = A.transpose(1, 2, 0)                                         # rows (s, r)
        qrow = -qs[:, None] + links[p + 1].charges()[None, :]
        qcol = links[ia].charges()[:, None] + cW[p][None, :]
    n1, n2, dim_a = Ar.shape
    nexp = compute_expansion(dim_a, n1 * n2, expansion_factor, max_expand, maxdim)
    if nexp <= 0:
        return local, []
    R, C = n1 * n2, S.shape[2] * S.shape[3]
    A2 = Ar.reshape(R, dim_a)
    S2 = S.reshape(R, C)
    for _ in range(north_pass):
        S2 = S2 - A2 @ (A2.T @ S2)
    qr_, qc_ = qrow.reshape(-1), qcol.reshape(-1)
    sectors = sorted(set(qr_.tolist()) & set(qc_.tolist()))
    ridx = [np.nonzero(qr_ == q)[0] for q in sectors]
    mats = [S2[np.ix_(ri, np.nonzero(qc_ == q)[0])] for q, ri in zip(sectors, ridx)]
    res, kept, _ = factorize_sectors(mats, False, cutoff=cutoff, maxdim=nexp, mindim=mindim, which_decomp="eigen")
    cols, ucharge = [], []
    for q, ri, r_ in zip(sectors, ridx, res):
        if r_ is None:
            continue
        U = np.zeros((R, r_[0].shape[1]))
        U[ri] = r_[0]
        cols.append(U)
        ucharge += [q] * U.shape[1]
    if not cols:
        return local, []
    U = np.concatenate(cols, axis=1)
    for _ in range(north_pass):
        U = U - A2 @ (A2.T @ U)
    if np.linalg.norm(U.T @ A2) > 1e-10:
        return local, []
    old = links[ia]
    qa = old.charges()
    ucharge = np.array(ucharge)
    new_sectors = sorted({q: 0 for q, _ in old.sectors} | {int(q): 0 for q in ucharge})
    dims = {q: int(np.sum(qa == q)) + int(np.sum(ucharge == q)) for q in new_sectors}
    mid = Link([(q, dims[q]) for q in new_sectors])
    mo = mid.offsets()
    K = sum(dims.values())
    Ax = np.zeros((R, K))
    for q in new_sectors:
        oc = np.nonzero(qa == q)[0]
        uc = np.nonzero(ucharge == q)[0]
        Ax[:, mo[q]:mo[q] + len(oc)] = A2[:, oc]
        Ax[:, mo[q] + len(oc):mo[q] + len(oc) + len(uc)] = U[:, uc]
    expander = Ax.T @ A2                                                  # [ax, a]
    links[ia] = mid
    if left:
        mps[p] = Ax.reshape(n1, n2, K)
        mps[v] = np.tensordot(expander, mps[v], axes=([1], [0]))
        local = np.tensordot(expander, local, axes=([1], [0]))
    else:
        mps[p] = Ax.reshape(n1, n2, K).transpose(2, 0, 1).copy()
        mps[v] = np.tensordot(mps[v], expander, axes=([2], [1]))
        local = np.tensordot(local, expander, axes=([local.ndim - 1], [1]))
    return local, [p, v]


def qn_qr_toward(mps, links, qs, x, y):
    """One step of the orthogonalize walk (extracter.jl:6) on a QN chain, moving the
    gauge centre from site x to its neighbour y: ITensors qr of a QN tensor is
    block-wise, each charge of the bond between x and y decomposed on its own
    (linalg.qr_pos gauge, thin: min(rows, cols) vectors per charge); site x
    becomes the isometry, y absorbs R."""
    A = mps[x]
    dl, d, dr = A.shape
    if y == x + 1:
        M = A.reshape(dl * d, dr)                                        # rows (l, s), columns r
        qrow = (links[x].charges()[:, None] + qs[None, :]).reshape(-1)
        bond = x + 1
    else:
        M = A.reshape(dl, d * dr).T                                      # rows (s, r), columns l
        qrow = (-qs[:, None] + links[x + 1].charges()[None, :]).reshape(-1)
        bond = x
    qcol = links[bond].charges()
    sec, Qs, Rs = [], [], []
    for q, _ in links[bond].sectors:
        ri, ci = np.nonzero(qrow == q)[0], np.nonzero(qcol == q)[0]
        if len(ri) == 0:
            continue
        Qq, Rq = linalg.qr_pos(M[np.ix_(ri, ci)])
        sec.append((int(q), Qq.shape[1]))
        Qs.append((ri, Qq))
        Rs.append((ci, Rq))
    K = sum(k for _, k in sec)
    Q = np.zeros((M.shape[0], K))
    R = np.zeros((K, M.shape[1]))
    o = 0
    for (ri, Qq), (ci, Rq) in zip(Qs, Rs):
        k = Qq.shape[1]
        Q[ri, o:o + k] = Qq
        R[o:o + k, ci] = Rq
        o += k
    links[bond] = Link(sec)
    if y == x + 1:
        mps[x] = Q.reshape(dl, d, K)
        mps[y] = np.tensordot(R, mps[y], axes=([1], [0]))
    else:
        mps[x] = Q.T.reshape(K, d, dr).copy()
        mps[y] = np.tensordot(mps[y], R.T, axes=([2], [0]))


@dataclass
class QNResult:
    energy: float
    mps: list
    links: list
    log: list = field(default_factory=list)


def qn_dmrg_chain(W, states, d, *, nsweeps, maxdim=None, cutoff=0.0, mindim=1, krylovdim=3, tol=1e-14,
                  which_decomp=None, subspace_algorithm=None, expansion_factor=1.5, max_expand=None, north_pass=1,
                  nsites=2):
    """QN DMRG from a product state (basis state states[i] on site i) over the
    Euler-tour plan from the last site (edges for nsites=2, vertices for
    nsites=1), with the densitymatrix subspace expansion in the extracter when
    asked (extracter.jl:10-12, the operator's previous position, the inserter's
    truncation); log rows (region, eigenvalue, kept, truncerr, new link sectors),
    kept and truncerr None for 1-site regions (inserter.jl stores the tensor)."""
    if nsites == 1:
        return _qn_dmrg_1site(W, states, d, nsweeps=nsweeps, maxdim=maxdim, cutoff=cutoff, mindim=mindim,
                              krylovdim=krylovdim, tol=tol, subspace_algorithm=subspace_algorithm,
                              expansion_factor=expansion_factor, max_expand=max_expand, north_pass=north_pass)
    N = len(W)
    qs = spin_charges(d)
    links = [Link([(0, 1)])]
    q = 0
    mps = []
    for s in states:
        q += qs[s]
        links.append(Link([(int(q), 1)]))
        a = np.zeros((1, d, 1))
        a[0, s, 0] = 1.0
        mps.append(a)
    if subspace_algorithm not in (None, "densitymatrix"):
        raise NotImplementedError("Subspace expansion (subspace_expand!) not defined for requested combination "
                                  "of subspace_algorithm and problem types")
    cW = mpo_link_charges(W, qs)
    envs = ChainEnvs(W)
    log, E = [], None
    pos = list(range(N))
    for sweep in range(1, nsweeps + 1):
        cut = get_or_last(cutoff, sweep)
        mxd = get_or_last(maxdim, sweep) if maxdim is not None else None
        mnd = get_or_last(mindim, sweep)
        for (u, v) in euler_sweep_chain(N, N - 1):
            lo, hi = min(u, v), max(u, v)
            th = theta2(mps[lo], mps[hi])
            if subspace_algorithm == "densitymatrix":
                th, changed = densitymatrix_expand_qn(mps, links, qs, cW, W, envs, pos, (lo, hi), th, cutoff=cut,
                                                      maxdim=mxd, mindim=mnd, expansion_factor=expansion_factor,
                                                      max_expand=max_expand, north_pass=north_pass)
                for c in changed:
                    envs.site_changed(c)
            pos = [lo, hi]
            L = envs.get(mps, lo - 1, lo)
            R = envs.get(mps, hi + 1, hi)
            shape = th.shape
            E, x, _ = linalg.lanczos_eigsolve(lambda y: heff2_apply(L, W[lo], W[hi], R, y.reshape(shape)).reshape(-1),
                                              th.reshape(-1), krylovdim=krylovdim, tol=tol)
            left, right, mid, te = qn_split(x.reshape(shape), links[lo], qs, links[hi + 1], u != lo, cutoff=cut,
                                            maxdim=mxd, mindim=mnd, which_decomp=which_decomp)
            mps[lo], mps[hi] = left, right
            links[lo + 1] = mid
            envs.site_changed(lo)
            envs.site_changed(hi)
            log.append(((u, v), E, sum(k for _, k in mid.sectors), te, list(mid.sectors)))
    return QNResult(E, mps, links, log)


def _product_chain(states, d):
    qs = spin_charges(d)
    links = [Link([(0, 1)])]
    q = 0
    mps = []
    for s in states:
        q += qs[s]
        links.append(Link([(int(q), 1)]))
        a = np.zeros((1, d, 1))
        a[0, s, 0] = 1.0
        mps.append(a)
    return mps, links


def _qn_dmrg_1site(W, states, d, *, nsweeps, maxdim, cutoff, mindim, krylovdim, tol, subspace_algorithm,
                   expansion_factor, max_expand, north_pass):
    """1-site QN DMRG: per region the QR walk brings the gauge centre to the site,
    the expansion (if asked) grows the link to the previous position, Lanczos on
    the 1-site H_eff, and the site stores the Ritz vector (inserter.jl:17-19)."""
    from .chain import heff1_apply
    from .plans import chain_graph, euler_tour_vertices
    N = len(W)
    qs = spin_charges(d)
    if subspace_algorithm not in (None, "densitymatrix"):
        raise NotImplementedError("Subspace expansion (subspace_expand!) not defined for requested combination "
                                  "of subspace_algorithm and problem types")
    cW = mpo_link_charges(W, qs)
    mps, links = _product_chain(states, d)
    envs = ChainEnvs(W)
    log, E = [], None
    pos = list(range(N))
    centre = N - 1                       # a product state is orthonormal about any site
    plan = euler_tour_vertices(*chain_graph(N), N - 1)
    for sweep in range(1, nsweeps + 1):
        cut = get_or_last(cutoff, sweep)
        mxd = get_or_last(maxdim, sweep) if maxdim is not None else None
        mnd = get_or_last(mindim, sweep)
        for v in plan:
            while centre != v:
                nxt = centre + (1 if v > centre else -1)
                qn_qr_toward(mps, links, qs, centre, nxt)
                envs.site_changed(centre)
                envs.site_changed(nxt)
                centre = nxt
            th = mps[v]
            if subspace_algorithm == "densitymatrix":
                th, changed = densitymatrix_expand_qn(mps, links, qs, cW, W, envs, pos, (v, v), th, cutoff=cut,
                                                      maxdim=mxd, mindim=mnd, expansion_factor=expansion_factor,
                                                      max_expand=max_expand, north_pass=north_pass)
                for c in changed:
                    envs.site_changed(c)
            pos = [v]
            L = envs.get(mps, v - 1, v)
            R = envs.get(mps, v + 1, v)
            shape = th.shape
            E, x, _ = linalg.lanczos_eigsolve(lambda y: heff1_apply(L, W[v], R, y.reshape(shape)).reshape(-1),
                                              th.reshape(-1), krylovdim=krylovdim, tol=tol)
            mps[v] = x.reshape(shape)
            envs.site_changed(v)
            log.append(((v,), E, None, None, None))
    return QNResult(E, mps, links, log)
