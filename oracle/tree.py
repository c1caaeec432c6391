"""Tree tensor networks on the CPU (test infrastructure only).

Restates the tree form of the sweep: ITensorNetworks' ``ttn``/``ProjTTN`` on a
tree with regions of one or two vertices (operator_map.jl:3-10 with up to four
incident environments, extracter.jl:3-17, inserter.jl:3-33), the DMRG driver
(eigsolve.jl:14-76) over the Euler-tour plan and the TDVP driver (applyexp.jl)
over the post-order plans.

Layout: vertex tensor A[l_n1, ..., l_nk, s] with one link per neighbour in
adjacency order (edge-list order, euler_tour.jl:4-11), MPO W[w_n1, ..., w_nk, s',
s], environment of the directed edge p -> q E[bra, mpo, ket].  Contractions use
numpy tensordot over named indices; the effective Hamiltonian contracts every
incident environment first and the MPO tensors last (a different order from the
device program, so the two are independent restatements).

Pinned by (tests/test_oracle_tree.py): the dense H of an MPO equal to the
brute-force sum of Kronecker products; tree DMRG against exact diagonalisation
(the reference's own build_tree(3,3) test); tree TDVP equal to expm(-iHt) at full
bond rank.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import linalg
from .plans import compute_adjacencies, euler_tour_edges, euler_tour_vertices, get_or_last, tdvp_regions
from .tdvp import krylov_expm, runge_kutta_4


def default_root(vertices, edges):
    """Last leaf in vertex order (the recalled NamedGraphs rule)."""
    adj = compute_adjacencies(vertices, edges)
    leaves = [v for v in vertices if len(adj[v]) <= 1]
    return leaves[-1] if leaves else vertices[-1]


class Tree:
    def __init__(self, vertices, edges):
        self.vertices = list(vertices)
        self.edges = [tuple(e) for e in edges]
        self.idx = {v: i for i, v in enumerate(self.vertices)}
        self.adj = compute_adjacencies(self.vertices, self.edges)

    def key(self, u, v):
        return (u, v) if self.idx[u] < self.idx[v] else (v, u)

    def labels(self, v):
        return [("k", self.key(v, n)) for n in self.adj[v]] + [("s", v)]

    def wlabels(self, v):
        return [("w", self.key(v, n)) for n in self.adj[v]] + [("t", v), ("s", v)]

    def side(self, u, v):
        seen, stack = {u}, [u]
        while stack:
            x = stack.pop()
            for y in self.adj[x]:
                if y not in seen and not (x == u and y == v):
                    seen.add(y)
                    stack.append(y)
        return seen

    def path(self, a, b):
        prev, stack = {a: None}, [a]
        while stack:
            x = stack.pop()
            for y in self.adj[x]:
                if y not in prev:
                    prev[y] = x
                    stack.append(y)
        out = [b]
        while out[-1] != a:
            out.append(prev[out[-1]])
        return out[::-1]


def _bra(lab):
    return ({"k": "b", "s": "t"}.get(lab[0], lab[0]),) + tuple(lab[1:])


def ltdot(A, la, B, lb, conj_b=False):
    common = [x for x in la if x in lb]
    out = np.tensordot(A, B.conj() if conj_b else B, axes=([la.index(x) for x in common], [lb.index(x) for x in common]))
    return out, [x for x in la if x not in common] + [x for x in lb if x not in common]


def lperm(A, la, target):
    return np.transpose(A, [la.index(x) for x in target])


# ------------------------------------------------------------------ dense forms
def ttn_to_dense(T: Tree, arrays):
    """State vector with site order = vertex order (small trees only)."""
    X, lx = arrays[T.vertices[0]], T.labels(T.vertices[0])
    for v in T.vertices[1:]:
        X, lx = ltdot(X, lx, arrays[v], T.labels(v))
    return lperm(X, lx, [("s", v) for v in T.vertices]).reshape(-1)


def mpo_to_dense(T: Tree, W):
    X, lx = W[T.vertices[0]], T.wlabels(T.vertices[0])
    for v in T.vertices[1:]:
        X, lx = ltdot(X, lx, W[v], T.wlabels(v))
    n = len(T.vertices)
    X = lperm(X, lx, [("t", v) for v in T.vertices] + [("s", v) for v in T.vertices])
    dim = int(np.prod(X.shape[:n]))
    return X.reshape(dim, dim)


def dense_hamiltonian(T: Tree, d, onsite=None, bond_terms=None):
    """Brute force sum of Kronecker products: sum_v onsite(v) + sum over edges (a, b)
    of the edge list of sum c O_a P_b (bond_terms: list, dict edge -> list, or
    callable edge -> list; onsite: matrix, dict, or None)."""
    n = len(T.vertices)
    dim = d ** n
    H = np.zeros((dim, dim))

    def embed(ops):
        M = np.ones((1, 1))
        for v in T.vertices:
            M = np.kron(M, ops.get(v, np.eye(d)))
        return M

    for v in T.vertices:
        h = None if onsite is None else (onsite.get(v) if isinstance(onsite, dict) else onsite)
        if h is not None:
            H += embed({v: np.asarray(h)})
    for (a, b) in T.edges:
        if bond_terms is None:
            terms = []
        elif callable(bond_terms):
            terms = bond_terms((a, b))
        elif isinstance(bond_terms, dict):
            terms = bond_terms.get((a, b), bond_terms.get((b, a), []))
        else:
            terms = bond_terms
        for c, O, P in terms:
            H += c * embed({a: np.asarray(O), b: np.asarray(P)})
    return H


# ------------------------------------------------------------------ sweep pieces
class TreeEnvs:
    """Environments per directed edge (u, v) = u's side seen from v, rebuilt when a
    vertex on that side changed (stamps)."""

    def __init__(self, T: Tree, W):
        self.T, self.W = T, W
        self.cache = {}
        self.clock = 0
        self.stamp = {v: 0 for v in T.vertices}
        self.sides = {}

    def touch(self, v):
        self.clock += 1
        self.stamp[v] = self.clock

    def side(self, u, v):
        if (u, v) not in self.sides:
            self.sides[(u, v)] = self.T.side(u, v)
        return self.sides[(u, v)]

    def get(self, psi, u, v):
        ent = self.cache.get((u, v))
        if ent is not None and all(self.stamp[x] <= ent[1] for x in self.side(u, v)):
            return ent[0]
        E = self.build(psi, u, v, psi[u])
        self.cache[(u, v)] = (E, self.clock)
        return E

    def build(self, psi, u, v, Au):
        T = self.T
        e = T.key(u, v)
        X, lx = Au, T.labels(u)
        for n in T.adj[u]:
            if n != v:
                en = T.key(n, u)
                X, lx = ltdot(X, lx, self.get(psi, n, u), [("b", en), ("w", en), ("k", en)])
        W = self.W[u].astype(np.result_type(self.W[u], Au))
        X, lx = ltdot(X, lx, W, T.wlabels(u))
        X, lx = ltdot(X, lx, Au, [_bra(x) for x in T.labels(u)], conj_b=True)
        return lperm(X, lx, [("b", e), ("w", e), ("k", e)])


def region_labels(T: Tree, region):
    if len(region) == 1:
        return T.labels(region[0])
    u, v = region
    e = T.key(u, v)
    return [x for x in T.labels(u) if x != ("k", e)] + [x for x in T.labels(v) if x != ("k", e)]


def heff_apply(T: Tree, W, envs_in, region, x):
    """Effective Hamiltonian of a region: every incident environment, then the MPO
    tensors of the region (envs_in: dict (n, r) -> environment)."""
    xl = region_labels(T, region)
    X, lx = x, list(xl)
    for (n, r), E in envs_in.items():
        e = T.key(n, r)
        X, lx = ltdot(X, lx, E, [("b", e), ("w", e), ("k", e)])
    for r in region:
        X, lx = ltdot(X, lx, W[r].astype(np.result_type(W[r], x)), T.wlabels(r))
    return lperm(X, lx, [_bra(y) for y in xl])


def _qr_toward(T, psi, x, p):
    e = T.key(x, p)
    lx = T.labels(x)
    rows = [y for y in lx if y != ("k", e)]
    A = lperm(psi[x], lx, rows + [("k", e)])
    shp = A.shape
    Q, R = linalg.qr_pos(A.reshape(-1, shp[-1]))
    psi[x] = lperm(Q.reshape(shp[:-1] + (Q.shape[1],)), rows + [("k", e)], lx)
    B, lb = ltdot(R, [("n", e), ("k", e)], psi[p], T.labels(p))
    psi[p] = lperm(B, [("k", e) if y == ("n", e) else y for y in lb], T.labels(p))


def orthogonalize(T: Tree, psi, ortho, region):
    """Returns the new ortho region; modifies psi (dict) in place."""
    rs = set(region)
    if ortho is not None and set(ortho) <= rs:
        return list(region)
    if ortho is None:
        parent, order, frontier, seen = {}, [], list(region), set(region)
        while frontier:
            nxt = []
            for x in frontier:
                for y in T.adj[x]:
                    if y not in seen:
                        seen.add(y)
                        parent[y] = x
                        order.append(y)
                        nxt.append(y)
            frontier = nxt
        for x in reversed(order):
            _qr_toward(T, psi, x, parent[x])
        return list(region)
    target = min(region, key=lambda r: min(len(T.path(c, r)) for c in ortho))
    c0 = min(ortho, key=lambda c: len(T.path(c, target)))
    for c in ortho:
        if c != c0:
            p = T.path(c, c0)
            for a, b in zip(p[:-1], p[1:]):
                _qr_toward(T, psi, a, b)
    p = T.path(c0, target)
    for a, b in zip(p[:-1], p[1:]):
        if a in rs:
            break
        _qr_toward(T, psi, a, b)
    return list(region)


@dataclass
class TreeResult:
    energy: float
    psi: dict
    log: list = field(default_factory=list)
    current_time: complex = 0.0


def compute_expansion(current_dim, basis_size, expansion_factor=1.5, max_expand=None, maxdim=None):
    """subspace.jl:28-48."""
    n = int(np.ceil(expansion_factor * current_dim))
    if max_expand is not None:
        n = min(max_expand, n)
    n = min(basis_size - current_dim, n)
    if maxdim is not None:
        n = min(maxdim - current_dim, n)
    return max(0, n)


def densitymatrix_expand(T, W, envs, psi, prev_pos, region, local, local_labels, *, cutoff, maxdim, mindim,
                         expansion_factor=1.5, max_expand=None, north_pass=1):
    """subspace/densitymatrix.jl:53-122: enlarge the link between the operator's
    previous position p and the region by the dominant eigenvectors of
    rho = S S^dag, S = (1 - P_A) (environments of p off the region . W_p . A),
    taken with the eigen-route truncation (cutoff, maxdim = the expansion size);
    A <- [A | U], the next vertex and the local tensor padded by A^dag [A | U].
    Returns (local, changed vertices)."""
    prev = [v for v in prev_pos if v not in region]
    if len(prev) != 1:
        return local, []
    p = prev[0]
    nxt = [v for v in region if v in T.adj[p]]
    if not nxt:
        return local, []
    v = nxt[0]
    A, la = psi[p], T.labels(p)
    e = T.key(p, v)
    a = ("k", e)
    ia = la.index(a)
    basis = int(np.prod([n for i, n in enumerate(A.shape) if i != ia]))
    nexp = compute_expansion(A.shape[ia], basis, expansion_factor, max_expand, maxdim)
    if nexp <= 0:
        return local, []
    S, ls = A, list(la)
    for n in T.adj[p]:
        if n not in region:
            en = T.key(n, p)
            S, ls = ltdot(S, ls, envs.get(psi, n, p), [("b", en), ("w", en), ("k", en)])
    S, ls = ltdot(S, ls, W[p].astype(np.result_type(W[p], A)), T.wlabels(p))
    lab = [_bra(x) for x in la]
    for _ in range(north_pass):
        X, lx = ltdot(A.conj(), lab, S, ls)
        Y, ly = ltdot(A, lab, X, lx)
        S = S - lperm(Y, ly, ls)
    rows = [x for x in ls if x[0] in ("b", "t")]
    cols = [x for x in ls if x not in rows]
    M = lperm(S, ls, rows + cols)
    nr = int(np.prod(M.shape[:len(rows)]))
    U, _, _ = linalg.factorize(M.reshape(nr, -1), cutoff=cutoff, maxdim=nexp, mindim=mindim, which_decomp="eigen")
    U = U.reshape(M.shape[:len(rows)] + (U.shape[1],))
    lu = [(("k",) + x[1:]) if x[0] == "b" else (("s",) + x[1:]) for x in rows] + [("x", e)]
    laa = [_bra(x) if x == a else x for x in la]
    for _ in range(north_pass):
        X, lx = ltdot(A.conj(), laa, U, lu)
        Y, ly = ltdot(A, laa, X, lx)
        U = U - lperm(Y, ly, lu)
    X, _ = ltdot(U.conj(), lu, A, la)
    if np.linalg.norm(X) > 1e-10:
        return local, []
    lo = [x for x in la if x != a]
    Ap = lperm(A, la, lo + [a])
    Up = lperm(U, lu, lo + [("x", e)])
    Ax = np.concatenate([Ap, Up], axis=-1)                  # directsum(A => a, U => new)
    expander = np.tensordot(Ax.conj(), Ap, axes=(list(range(len(lo))), list(range(len(lo)))))   # [ax, a]
    psi[p] = lperm(Ax, lo + [a], la)
    Cn, lc = ltdot(expander, [("xa", e), a], psi[v], T.labels(v))
    psi[v] = lperm(Cn, [a if x == ("xa", e) else x for x in lc], T.labels(v))
    Ln, ll = ltdot(expander, [("xa", e), a], local, local_labels)
    return lperm(Ln, [a if x == ("xa", e) else x for x in ll], local_labels), [p, v]


def tree_dmrg(vertices, edges, W, psi0, *, nsweeps, nsites=2, maxdim=None, cutoff=0.0, mindim=1, krylovdim=3,
              tol=1e-14, root_vertex=None, which_decomp=None, subspace_algorithm=None, expansion_factor=1.5,
              max_expand=None, north_pass=1, expand_trunc=None):
    """DMRG on a tree (eigsolve.jl:14-76) over the Euler-tour plan (euler_plans.jl:4-13;
    edges for nsites=2, vertices for nsites=1), with the densitymatrix subspace
    expansion in the extracter when asked (extracter.jl:10-12, using the
    operator's previous position); log rows (region, eigenvalue, kept, truncerr).
    ``expand_trunc`` (dict cutoff/maxdim/mindim) is the extracter's truncation;
    by default the inserter's."""
    T = Tree(vertices, edges)
    cplx = any(np.iscomplexobj(a) for a in psi0.values())
    psi = {v: np.array(a, dtype=np.complex128 if cplx else np.float64) for v, a in psi0.items()}
    envs = TreeEnvs(T, W)
    root = default_root(T.vertices, T.edges) if root_vertex is None else root_vertex
    ortho = None
    pos = list(T.vertices)
    log, E = [], None
    if nsites == 2:
        plan = [list(r) for r in euler_tour_edges(T.vertices, T.edges, root)]
    else:
        plan = [[x] for x in euler_tour_vertices(T.vertices, T.edges, root)]
    for sweep in range(1, nsweeps + 1):
        cut = get_or_last(cutoff, sweep)
        mxd = get_or_last(maxdim, sweep) if maxdim is not None else None
        mnd = get_or_last(mindim, sweep)
        et = expand_trunc or {}
        ecut = get_or_last(et.get("cutoff", cutoff), sweep)
        emx = et.get("maxdim", maxdim)
        emx = get_or_last(emx, sweep) if emx is not None else None
        emn = get_or_last(et.get("mindim", mindim), sweep)
        for region in plan:
            before = {x: psi[x] for x in T.vertices}
            ortho = orthogonalize(T, psi, ortho, region)
            for x in T.vertices:
                if psi[x] is not before[x]:
                    envs.touch(x)
            if len(region) == 2:
                u, v = region
                th, _ = ltdot(psi[u], T.labels(u), psi[v], T.labels(v))
            else:
                th = psi[region[0]]
            if subspace_algorithm == "densitymatrix":
                th, changed = densitymatrix_expand(T, W, envs, psi, pos, region, th, region_labels(T, region),
                                                   cutoff=ecut, maxdim=emx, mindim=emn,
                                                   expansion_factor=expansion_factor, max_expand=max_expand,
                                                   north_pass=north_pass)
                for x in changed:
                    envs.touch(x)
            pos = list(region)
            envs_in = {(n, r): envs.get(psi, n, r) for r in region for n in T.adj[r] if n not in region}
            shape = th.shape
            E, x, _ = linalg.lanczos_eigsolve(
                lambda y: heff_apply(T, W, envs_in, region, y.reshape(shape)).reshape(-1), th.reshape(-1),
                krylovdim=krylovdim, tol=tol)
            x = x.reshape(shape)
            if len(region) == 2:
                u, v = region
                e = T.key(u, v)
                lu = [y for y in T.labels(u) if y != ("k", e)]
                lv = [y for y in T.labels(v) if y != ("k", e)]
                nu = int(np.prod([shape[i] for i in range(len(lu))]))
                U, C, info = linalg.factorize(x.reshape(nu, -1), cutoff=cut, maxdim=mxd, mindim=mnd,
                                              which_decomp=which_decomp)
                psi[u] = lperm(U.reshape(shape[:len(lu)] + (U.shape[1],)), lu + [("k", e)], T.labels(u))
                psi[v] = lperm(C.reshape((C.shape[0],) + shape[len(lu):]), [("k", e)] + lv, T.labels(v))
                envs.touch(u)
                log.append((tuple(region), E, info.kept, info.truncerr))
            else:
                v = region[0]
                psi[v] = x
                log.append((tuple(region), E, None, None))
            envs.touch(v)
            ortho = [v]
    return TreeResult(E, psi, log)


def tree_applyexp(vertices, edges, W, psi0, exponents, *, nsites=2, tdvp_order=4, solver="exponentiate",
                  maxdim=None, cutoff=0.0, mindim=1, normalize=False, krylovdim=30, tol=1e-12, root_vertex=None):
    """applyexp (applyexp.jl:18-89) on a tree: one sweep per exponent increment,
    tdvp_regions over the post-order plan, the local solver per region, for
    nsites=1 the backward evolution of the bond to the next region (first edge of
    the path), factorize with U on first(region)."""
    T = Tree(vertices, edges)
    cplx = any(complex(z).imag != 0 for z in exponents) or any(np.iscomplexobj(a) for a in psi0.values())
    dt = np.complex128 if cplx else np.float64
    psi = {v: np.array(a, dtype=dt) for v, a in psi0.items()}
    envs = TreeEnvs(T, W)
    root = default_root(T.vertices, T.edges) if root_vertex is None else root_vertex
    pts = [0.0] + list(exponents)
    steps = [pts[i + 1] - pts[i] for i in range(len(pts) - 1)][1:]
    if not cplx:
        steps = [float(np.real(z)) for z in steps]
    ortho = None
    now = 0.0

    def solve(apply, tau, x):
        if solver == "exponentiate":
            return krylov_expm(apply, tau, x, krylovdim=krylovdim, tol=tol)[0]
        return runge_kutta_4(apply, tau, x)

    for sweep, dtau in enumerate(steps, start=1):
        cut = get_or_last(cutoff, sweep)
        mxd = get_or_last(maxdim, sweep) if maxdim is not None else None
        mnd = get_or_last(mindim, sweep)
        plan = tdvp_regions(T.vertices, T.edges, root, dtau, tdvp_order, nsites)
        for idx, (region, tau) in enumerate(plan):
            region = list(region)
            before = {x: psi[x] for x in T.vertices}
            ortho = orthogonalize(T, psi, ortho, region)
            for x in T.vertices:
                if psi[x] is not before[x]:
                    envs.touch(x)
            envs_in = {(n, r): envs.get(psi, n, r) for r in region for n in T.adj[r] if n not in region}
            if len(region) == 2:
                u, v = region
                th, _ = ltdot(psi[u], T.labels(u), psi[v], T.labels(v))
            else:
                th = psi[region[0]]
            shape = th.shape
            x = solve(lambda y: heff_apply(T, W, envs_in, region, y.reshape(shape)).reshape(-1), tau,
                      th.reshape(-1)).reshape(shape)
            if len(region) == 2:
                u, v = region
                e = T.key(u, v)
                lu = [y for y in T.labels(u) if y != ("k", e)]
                lv = [y for y in T.labels(v) if y != ("k", e)]
                nu = int(np.prod([shape[i] for i in range(len(lu))]))
                U, C, _ = linalg.factorize(x.reshape(nu, -1), cutoff=cut, maxdim=mxd, mindim=mnd)
                psi[u] = lperm(U.reshape(shape[:len(lu)] + (U.shape[1],)), lu + [("k", e)], T.labels(u))
                psi[v] = lperm(C.reshape((C.shape[0],) + shape[len(lu):]), [("k", e)] + lv, T.labels(v))
                envs.touch(u)
            else:
                (v,) = region
                if nsites == 1 and idx + 1 < len(plan) and list(plan[idx + 1][0]) != region:
                    x = _back_evolve(T, W, envs, psi, v, plan[idx + 1][0][0], x, tau, solve)
                psi[v] = x
            if normalize:
                psi[v] = psi[v] / np.linalg.norm(psi[v])
            envs.touch(v)
            ortho = [v]
            now += tau
    return TreeResult(None, psi, [], now)


def _back_evolve(T, W, envs, psi, v, target, x, tau, solve):
    w = T.path(v, target)[1]
    e = T.key(v, w)
    lv = T.labels(v)
    rows = [y for y in lv if y != ("k", e)]
    A = lperm(x, lv, rows + [("k", e)])
    shp = A.shape
    Q, R = linalg.qr_pos(A.reshape(-1, shp[-1]))
    Qc = lperm(Q.reshape(shp[:-1] + (Q.shape[1],)), rows + [("k", e)], lv)
    Eq = envs.build(psi, v, w, Qc)                      # [m', w, m]
    Ew = envs.get(psi, w, v)                            # [b', w, b]

    def bond(y):
        Y = y.reshape(R.shape)
        Z = np.tensordot(Eq, Y, axes=([2], [0]))        # [m', w, b]
        return np.tensordot(Z, Ew, axes=([1, 2], [1, 2])).reshape(-1)   # [m', b']

    Rt = solve(bond, -tau, R.reshape(-1)).reshape(R.shape)
    out, lo = ltdot(Qc, [("k", ("q", e)) if y == ("k", e) else y for y in lv], Rt, [("k", ("q", e)), ("k", e)])
    return lperm(out, lo, lv)


def expectation(T: Tree, W, psi):
    """<psi|H|psi> / <psi|psi> by dense-free contraction (all environments toward
    the last vertex)."""
    r = T.vertices[-1]
    envs = TreeEnvs(T, W)
    envs_in = {(n, r): envs.get(psi, n, r) for n in T.adj[r]}
    y = heff_apply(T, W, envs_in, [r], psi[r])
    num = np.vdot(psi[r], y)
    # norm: all environments with the identity operator
    Wid = {}
    for v in T.vertices:
        k = len(T.adj[v])
        d = psi[v].shape[-1]
        Wid[v] = np.eye(d).reshape([1] * k + [d, d])
    envs1 = TreeEnvs(T, Wid)
    n_in = {(n, r): envs1.get(psi, n, r) for n in T.adj[r]}
    den = np.vdot(psi[r], heff_apply(T, Wid, n_in, [r], psi[r]))
    return (num / den).real


def inner(T: Tree, a, b):
    """<a|b>: leaves first toward the last vertex, each vertex contracted with the
    transfer matrices of its children (intermediates carry one link pair)."""
    root = T.vertices[-1]
    parent, order, stack = {root: None}, [], [root]
    while stack:
        x = stack.pop()
        order.append(x)
        for y in T.adj[x]:
            if y not in parent:
                parent[y] = x
                stack.append(y)
    up = {}
    for v in reversed(order):
        lb = [("b", T.key(v, n)) for n in T.adj[v]] + [("t", v)]
        lk = [("k", T.key(v, n)) for n in T.adj[v]] + [("t", v)]
        X, lx = b[v], lk
        for n in T.adj[v]:
            if n != parent[v]:
                e = T.key(n, v)
                X, lx = ltdot(X, lx, up[n], [("b", e), ("k", e)])
        X, lx = ltdot(X, lx, a[v].conj(), lb)
        if parent[v] is None:
            return X.reshape(-1)[0]
        e = T.key(v, parent[v])
        up[v] = lperm(X, lx, [("b", e), ("k", e)])


def local_expectations(T: Tree, psi, op):
    """<psi|op_v|psi> / <psi|psi> for every vertex, each by a full contraction of
    <psi| with op applied at v (no gauge assumption)."""
    den = inner(T, psi, psi).real
    out = {}
    for v in T.vertices:
        phi = dict(psi)
        phi[v] = np.tensordot(psi[v], np.asarray(op), axes=([psi[v].ndim - 1], [1]))
        out[v] = float(np.real(inner(T, psi, phi)) / den)
    return out
