"""Subspace expansion in the extracter (subspace/subspace.jl, subspace/densitymatrix.jl).

``subspace_expand`` dispatches on ``subspace_algorithm`` as the reference's
backend selection does: None does nothing, "densitymatrix" enlarges the link
between the operator's previous position and the new region by the dominant
eigenvectors of the density matrix of H applied to the previous vertex, projected
off that vertex's current span; any other name raises the reference's error.

All contractions are device tensordots over labelled indices, so chains (site
tensor A[l, s, r], boundary links of dimension 1) and trees (A[links..., s]) share
one implementation through a small label adapter; QN chains take the block-sparse
form in ``qn.densitymatrix_expand``.
"""
from __future__ import annotations

import math

import numpy as np

from . import kernels, qn
from .tree import TreeTTN, _bra, _permute_to, _tdot

DEFAULT_EXPANSION_FACTOR = 1.5          # subspace.jl:5
DEFAULT_MAX_EXPAND = None               # typemax(Int), subspace.jl:6


def compute_expansion(current_dim, basis_size, *, expansion_factor=DEFAULT_EXPANSION_FACTOR,
                      max_expand=DEFAULT_MAX_EXPAND, maxdim=None):
    """subspace.jl:28-48: the number of new link directions (added to current_dim)."""
    n = math.ceil(expansion_factor * current_dim)
    if max_expand is not None:
        n = min(max_expand, n)
    n = min(basis_size - current_dim, n)
    if maxdim is not None:
        n = min(maxdim - current_dim, n)
    return max(0, n)


# ----------------------------------------------------------------- label views
class _ChainView:
    """Labels of a chain state A[l, s, r] with the boundary links as edges
    (-1, 0) and (N-1, N), whose environments are the trivial ones."""

    def __init__(self, psi, P):
        self.psi, self.P, self.N = psi, P, len(psi)

    @staticmethod
    def key(a, b):
        return ("c", min(a, b), max(a, b))

    def neighbors(self, v):
        return [v - 1, v + 1]

    def adjacent(self, u, v):
        return abs(u - v) == 1

    def labels(self, v):
        return [("k", self.key(v - 1, v)), ("s", v), ("k", self.key(v, v + 1))]

    def mpo(self, v):
        W = self.P.operator().tensor(v, self.psi.dtype)
        return W, [("w", self.key(v - 1, v)), ("w", self.key(v, v + 1)), ("t", v), ("s", v)]

    def env(self, n, v):
        k = self.key(n, v)
        return self.P.environment(self.psi, (n, v)), [("b", k), ("w", k), ("k", k)]

    def local_labels(self, region):
        lo, hi = min(region), max(region)
        if len(region) == 1:
            return self.labels(lo)
        return [("k", self.key(lo - 1, lo)), ("s", lo), ("s", hi), ("k", self.key(hi, hi + 1))]


class _TreeView:
    def __init__(self, psi, P):
        self.psi, self.P, self.tg = psi, P, psi.tg

    def key(self, a, b):
        return self.tg.key(a, b)

    def neighbors(self, v):
        return list(self.tg.adj[v])

    def adjacent(self, u, v):
        return v in self.tg.adj[u]

    def labels(self, v):
        return self.tg.site_labels(v)

    def mpo(self, v):
        return self.P.operator().tensor(v, self.psi.dtype), self.tg.mpo_labels(v)

    def env(self, n, v):
        k = self.key(n, v)
        return self.P.environment(self.psi, (n, v)), [("b", k), ("w", k), ("k", k)]

    def local_labels(self, region):
        from .tree import region_labels
        return region_labels(self.psi, list(region))


def _embed(ctx, n_from, n_total, offset, dtype):
    E = np.zeros((n_from, n_total), dtype=dtype)
    E[np.arange(n_from), offset + np.arange(n_from)] = 1.0
    return ctx.upload(E)


def densitymatrix_expand(problem, local_state, region, *, trunc, north_pass=1,
                         expansion_factor=DEFAULT_EXPANSION_FACTOR, max_expand=DEFAULT_MAX_EXPAND, **kws):
    """densitymatrix.jl:53-122 on the device.  Returns (problem, local_state)."""
    psi, P = problem.state, problem.operator
    if isinstance(psi, qn.QNMPS):
        return qn.densitymatrix_expand(problem, local_state, region, trunc=trunc, north_pass=north_pass,
                                       expansion_factor=expansion_factor, max_expand=max_expand)
    view = _TreeView(psi, P) if isinstance(psi, TreeTTN) else _ChainView(psi, P)
    region = list(region)
    prev = [v for v in P.sites() if v not in region]
    if len(prev) != 1:
        return problem, local_state
    p = prev[0]
    nxt = [v for v in region if view.adjacent(p, v)]
    if not nxt:
        return problem, local_state
    v = nxt[0]
    A = psi[p]
    la = view.labels(p)
    e = view.key(p, v)
    a = ("k", e)
    ia = la.index(a)
    dim_a = A.shape[ia]
    basis_size = int(np.prod([n for i, n in enumerate(A.shape) if i != ia]))
    nexp = compute_expansion(dim_a, basis_size, expansion_factor=expansion_factor, max_expand=max_expand,
                             maxdim=trunc.get("maxdim"))
    if nexp <= 0:
        return problem, local_state
    ctx = A.ctx
    # sqrt_rho = A . (environments of the previous position off the region) . W_p
    S, ls = A, list(la)
    for n in view.neighbors(p):
        if n in region:
            continue
        E, le = view.env(n, p)
        S, ls = _tdot(S, ls, E, le)
    W, lw = view.mpo(p)
    S, ls = _tdot(S, ls, W, lw)
    lA_bra = [_bra(x) for x in la]                          # prime(A)
    for _ in range(north_pass):                             # S <- S - A'(A'^dag S)
        X, lx = _tdot(A, lA_bra, S, ls, conj_a=True)
        Y, ly = _tdot(A, lA_bra, X, lx)
        S = S.copy().axpy_(-1.0, _permute_to(Y, ly, ls))
    rows = [x for x in ls if x[0] in ("b", "t")]            # primed: the other links and the site
    cols = [x for x in ls if x not in rows]                 # a and the MPO link toward v
    M = _permute_to(S, ls, rows + cols)
    U, _, _ = kernels.factorize(M, len(rows), False, cutoff=trunc.get("cutoff", 0.0), maxdim=nexp,
                                mindim=trunc.get("mindim", 1), which_decomp="eigen")
    new = ("x", e)
    lu = [(("k",) + x[1:]) if x[0] == "b" else (("s",) + x[1:]) for x in rows] + [new]   # noprime(rows) + new
    lA_a = [_bra(x) if x == a else x for x in la]           # prime(A, a)
    for _ in range(north_pass):                             # U <- U - A'(A'^dag U)
        X, lx = _tdot(A, lA_a, U, lu, conj_a=True)
        Y, ly = _tdot(A, lA_a, X, lx)
        U = U.copy().axpy_(-1.0, _permute_to(Y, ly, lu))
    X, _ = _tdot(U, lu, A, la, conj_a=True)
    if X.norm() > 1e-10:
        print(f"Warning: |U*A| = {X.norm():.3E} in subspace expansion")
        return problem, local_state
    k = U.shape[-1]
    # Ax = [A | U] along the link toward v
    lo = [x for x in la if x != a]
    Ap = _permute_to(A, la, lo + [a])
    Up = _permute_to(U, lu, lo + [new])
    xa = ("xa", e)
    Ax, lx1 = _tdot(Ap, lo + [a], _embed(ctx, dim_a, dim_a + k, 0, psi.dtype), [a, xa])
    Ux, _ = _tdot(Up, lo + [new], _embed(ctx, k, dim_a + k, dim_a, psi.dtype), [new, xa])
    Ax.axpy_(1.0, Ux)
    expander, le = _tdot(Ax, lx1, A, la, conj_a=True)     # [xa, a]
    Cn, lc = _tdot(expander, le, psi[v], view.labels(v))
    relabel = lambda labs: [a if x == xa else x for x in labs]     # noqa: E731
    psi[p] = _permute_to(Ax, relabel(lx1), la)
    psi[v] = _permute_to(Cn, relabel(lc), view.labels(v))
    ll = view.local_labels(region)
    Ln, lln = _tdot(expander, le, local_state, ll)
    local_state = _permute_to(Ln, relabel(lln), ll)
    problem.state = psi
    return problem, local_state


def subspace_expand(problem, local_state, region, *, subspace_algorithm=None, trunc, **kws):
    """subspace.jl:8-26: backend selection by name."""
    if subspace_algorithm is None:
        return problem, local_state
    if subspace_algorithm == "densitymatrix":
        return densitymatrix_expand(problem, local_state, region, trunc=trunc, **kws)
    raise NotImplementedError(
        "Subspace expansion (subspace_expand!) not defined for requested combination of subspace_algorithm "
        "and problem types")
