"""MPS/MPO tensor algebra and the 2-site DMRG region loop on a chain (test infrastructure only).

Layouts (row-major, last index fastest) -- the same as the device path:
  site tensor      A[l, s, r]          (chi_l, d, chi_r)
  MPO tensor       W[wl, wr, s', s]    (D_l, D_r, d, d); s' = output (bra), s = input (ket)
  environment      E[bra, mpo, ket]    (chi, D, chi) for every directed edge p->q,
                                        covering p's side of the chain
  two-site tensor  theta[l, s1, s2, r] in chain order whatever the region order
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from . import linalg
from .plans import euler_sweep_chain, get_or_last


# ---------------------------------------------------------------------------
# contractions (ITensorNetworks ProjTTN / ITensors contract restated)
# ---------------------------------------------------------------------------
def env_left(E, A, W):
    """E'[b', w', b] = sum conj(A[a', s', b']) E[a', w, a] W[w, w', s', s] A[a, s, b].

    The new environment of edge i->i+1 from that of i-1->i (ITensorNetworks
    ``make_environment`` via ``position``, extracter.jl:14).  Pinned by the
    brute-force <psi|H|psi> split test (tests/test_oracle_chain.py)."""
    X = np.tensordot(E, A, axes=([2], [0]))                 # [a', w, s, b]
    Y = np.tensordot(X, W, axes=([1, 2], [0, 3]))           # [a', b, w', s']
    return np.tensordot(A.conj(), Y, axes=([0, 1], [0, 3])).transpose(0, 2, 1).copy()


def env_right(E, A, W):
    """E'[a', w, a] = sum conj(A[a', s', b']) E[b', v, b] W[w, v, s', s] A[a, s, b].

    The new environment of edge i->i-1 from that of i+1->i (as env_left)."""
    X = np.tensordot(A.conj(), E, axes=([2], [0]))          # [a', s', v, b]
    Y = np.tensordot(X, W, axes=([1, 2], [2, 1]))           # [a', b, w, s]
    return np.tensordot(Y, A, axes=([1, 3], [2, 1])).copy()  # [a', w, a]


def theta2(A, B):
    """theta[l, s1, s2, r] = sum_m A[l, s1, m] B[m, s2, r] (extracter.jl:7)."""
    return np.tensordot(A, B, axes=([2], [0]))


def heff2_apply(L, W1, W2, R, theta):
    """Two-site effective Hamiltonian apply (operator_map.jl:3-10, ``optimal_map``):
    y[a', s1', s2', b'] = sum L[a', w, a] W1[w, u, s1', s1] W2[u, v, s2', s2]
                              R[b', v, b] theta[a, s1, s2, b]
    in the order ((L.theta).W1W2).R.  Pinned by the Kronecker-product matrix
    test and Hermiticity (tests/test_oracle_chain.py)."""
    T1 = np.tensordot(L, theta, axes=([2], [0]))            # [a', w, s1, s2, b]
    W12 = np.tensordot(W1, W2, axes=([1], [0]))             # [w, s1', s1, v, s2', s2]
    T3 = np.tensordot(T1, W12, axes=([1, 2, 3], [0, 2, 5]))  # [a', b, s1', v, s2']
    y = np.tensordot(T3, R, axes=([1, 3], [2, 1]))          # [a', s1', s2', b']
    return y


def heff1_apply(L, W, R, x):
    """One-site effective Hamiltonian apply (the 1-site regions of TDVP):
    y[a', s', b'] = sum L[a', w, a] W[w, v, s', s] R[b', v, b] x[a, s, b]."""
    T1 = np.tensordot(L, x, axes=([2], [0]))                # [a', w, s, b]
    T2 = np.tensordot(T1, W, axes=([1, 2], [0, 3]))         # [a', b, v, s']
    return np.tensordot(T2, R, axes=([1, 2], [2, 1]))       # [a', s', b']


def heff2_dense_matrix(L, W1, W2, R):
    """Dense matrix of the two-site effective Hamiltonian on vec(theta) (row-major),
    built by Kronecker products summed over MPO channels -- a brute-force
    restatement independent of the contraction order above."""
    D0, D1, D2 = W1.shape[0], W1.shape[1], W2.shape[1]
    cl, cr = L.shape[0], R.shape[0]
    d1, d2 = W1.shape[2], W2.shape[2]
    n = cl * d1 * d2 * cr
    H = np.zeros((n, n), dtype=np.result_type(L, W1, W2, R))
    for w in range(D0):
        for u in range(D1):
            for v in range(D2):
                H += np.kron(np.kron(np.kron(L[:, w, :], W1[w, u]), W2[u, v]), R[:, v, :])
    return H


# ---------------------------------------------------------------------------
# MPS utilities
# ---------------------------------------------------------------------------
def random_mps(N, d, chi, seed, dtype=np.float64):
    """Seeded random MPS: bond dims min(chi, d^k, d^(N-k)), entries i.i.d. N(0,1)
    from numpy PCG64(seed), drawn site by site in chain order (real part then
    imaginary part for complex types), then scaled to unit norm.  (The reference
    draws from Julia's RNG, which cannot be reproduced here; this generator is
    the shared input format of the oracle and the device path.)"""
    rng = np.random.Generator(np.random.PCG64(seed))
    dims = [1]
    for k in range(1, N):
        dims.append(int(min(chi, d ** k, d ** (N - k))))
    dims.append(1)
    mps = []
    for i in range(N):
        shape = (dims[i], d, dims[i + 1])
        a = rng.standard_normal(shape)
        if np.issubdtype(np.dtype(dtype), np.complexfloating):
            a = a + 1j * rng.standard_normal(shape)
        mps.append(a.astype(dtype))
    nrm = np.sqrt(abs(mps_inner(mps, mps)))
    mps[0] = mps[0] / nrm
    return mps


def product_mps(states, d, dtype=np.float64):
    """Product state MPS with site i in basis state states[i]."""
    mps = []
    for s in states:
        a = np.zeros((1, d, 1), dtype=dtype)
        a[0, s, 0] = 1.0
        mps.append(a)
    return mps


def mps_inner(A, B):
    """<A|B> by a left-to-right transfer contraction."""
    E = np.ones((1, 1), dtype=np.result_type(A[0], B[0]))
    for a, b in zip(A, B):
        X = np.tensordot(E, b, axes=([1], [0]))
        E = np.tensordot(a.conj(), X, axes=([0, 1], [0, 1]))
    return E[0, 0]


def mps_to_dense(mps):
    v = mps[0]
    for a in mps[1:]:
        v = np.tensordot(v, a, axes=([-1], [0]))
    return v.reshape(-1)


def mpo_to_dense(W):
    """Dense operator of an MPO (small N only)."""
    T = W[0]  # [1, D, s', s]
    T = T[0]  # [D, s', s]
    for w in W[1:]:
        T = np.tensordot(T, w, axes=([0], [0]))  # [..., D', s', s]
        nd = T.ndim
        # move the new MPO bond to the front
        T = np.moveaxis(T, nd - 3, 0)
    T = T[0]  # [s1', s1, s2', s2, ...]
    N = len(W)
    perm = [2 * i for i in range(N)] + [2 * i + 1 for i in range(N)]
    T = T.transpose(perm)
    dim = int(np.prod(T.shape[:N]))
    return T.reshape(dim, dim)


def mpo_expectation(W, mps):
    E = np.ones((1, 1, 1), dtype=mps[0].dtype)
    for w, a in zip(W, mps):
        E = env_left(E, a, w)
    return E[0, 0, 0]


def orthogonalize_chain(mps, lo, hi):
    """QR walk making sites < lo left-orthogonal and sites > hi right-orthogonal,
    pushing the R factors into the region [lo, hi] (ITensorNetworks
    ``orthogonalize(psi, region)`` as called at extracter.jl:6, on a chain whose
    orthogonality centre is unknown).  Bond dimensions shrink to the QR rank."""
    mps = [a.copy() for a in mps]
    for i in range(0, lo):
        a = mps[i]
        cl, d, cr = a.shape
        Q, Rm = linalg.qr_pos(a.reshape(cl * d, cr))
        mps[i] = Q.reshape(cl, d, Q.shape[1])
        mps[i + 1] = np.tensordot(Rm, mps[i + 1], axes=([1], [0]))
    for i in range(len(mps) - 1, hi, -1):
        a = mps[i]
        cl, d, cr = a.shape
        Q, Rm = linalg.qr_pos(a.reshape(cl, d * cr).T)       # A^T = Q R  ->  A = R^T Q^T
        mps[i] = Q.T.reshape(Q.shape[1], d, cr)
        mps[i - 1] = np.tensordot(mps[i - 1], Rm.T, axes=([2], [0]))
    return mps


# ---------------------------------------------------------------------------
# environments cache (ProjTTN on a chain)
# ---------------------------------------------------------------------------
class ChainEnvs:
    """Environments keyed by directed edge (p, q) = p's side seen from q.

    ``position(mps, lo, hi)`` makes the environments of the two edges incident
    to the region [lo, hi] current, building missing ones recursively from the
    boundary ones (ITensorNetworks ProjTTN ``position``/``make_environment``,
    extracter.jl:14).  ``site_changed(i)`` drops every cached environment whose
    side contains site i."""

    def __init__(self, W):
        self.W = W
        self.N = len(W)
        self.cache = {}
        self.builds = 0

    def one(self, dtype):
        return np.ones((1, 1, 1), dtype=dtype)

    def get(self, mps, p, q):
        if p < 0 or p >= self.N:
            return self.one(mps[0].dtype)
        key = (p, q)
        if key not in self.cache:
            if q == p + 1:
                prev = self.get(mps, p - 1, p)
                self.cache[key] = env_left(prev, mps[p], self.W[p])
            else:
                prev = self.get(mps, p + 1, p)
                self.cache[key] = env_right(prev, mps[p], self.W[p])
            self.builds += 1
        return self.cache[key]

    def site_changed(self, i):
        drop = [k for k in self.cache if (k[1] == k[0] + 1 and k[0] >= i) or
                (k[1] == k[0] - 1 and k[0] <= i)]
        for k in drop:
            del self.cache[k]


# ---------------------------------------------------------------------------
# 2-site DMRG on a chain: eigsolve.jl:45-76 + iterators.jl + extracter/updater/inserter
# ---------------------------------------------------------------------------
@dataclass
class RegionLog:
    sweep: int
    region: tuple
    energy: float
    kept: int
    truncerr: float
    route: str
    seconds: float = 0.0   # wall time of this region step (after any initial gauge walk)


@dataclass
class DMRGResult:
    energy: float
    mps: list
    log: list = field(default_factory=list)
    sweep_energies: list = field(default_factory=list)
    sweep_maxlinkdim: list = field(default_factory=list)
    seconds: float = 0.0


def dmrg_chain(W, mps0, *, nsweeps, maxdim=None, cutoff=0.0, mindim=1,
               krylovdim=3, maxiter=1, tol=1e-14, root_vertex=None,
               which_decomp=None, normalize=False, max_regions=None, max_seconds=None):
    """2-site DMRG on a chain, restating the region loop of
    sweep_solve.jl:12-40 -> iterators.jl:81-100 with
      plan      euler_sweep(nsites=2) from root_vertex (euler_plans.jl:4-13),
      extracter orthogonalize + theta + position (extracter.jl:3-17),
      updater   eigsolve_solver -> Lanczos (eigsolve.jl:14-28,
                local_solvers/eigsolve.jl:3-29),
      inserter  factorize with U on first(region), C on last(region)
                (inserter.jl:3-33),
    with per-sweep truncation parameters (truncation_parameters.jl:7-14).
    Returns the last region's Ritz value as the energy (eigsolve.jl:23,66).
    ``max_regions`` / ``max_seconds`` stop early (bounded timing samples only).
    """
    N = len(W)
    if root_vertex is None:
        root_vertex = N - 1
    mps = [a.copy() for a in mps0]
    envs = ChainEnvs(W)
    ortho = None  # unknown orthogonality centre
    energy = float("inf")
    res = DMRGResult(energy, mps)
    t0 = time.perf_counter()
    nreg = 0
    for sweep in range(1, nsweeps + 1):
        cut = get_or_last(cutoff, sweep)
        mxd = get_or_last(maxdim, sweep) if maxdim is not None else None
        mnd = get_or_last(mindim, sweep)
        for (u, v) in euler_sweep_chain(N, root_vertex):
            lo, hi = min(u, v), max(u, v)
            if ortho is None or not (lo <= ortho <= hi):
                mps = orthogonalize_chain(mps, lo, hi)
                for i in range(N):
                    envs.site_changed(i)
            t_region = time.perf_counter()
            th = theta2(mps[lo], mps[hi])
            L = envs.get(mps, lo - 1, lo)
            R = envs.get(mps, hi + 1, hi)
            shape = th.shape
            energy, x, _ = linalg.lanczos_eigsolve(
                lambda t: heff2_apply(L, W[lo], W[hi], R, t.reshape(shape)).reshape(-1),
                th.reshape(-1), krylovdim=krylovdim, maxiter=maxiter, tol=tol)
            th = x.reshape(shape)
            cl, d1, d2, cr = shape
            M = th.reshape(cl * d1, d2 * cr)
            if u == lo:   # moving right: U on lo
                U, C, info = linalg.factorize(M, cutoff=cut, maxdim=mxd, mindim=mnd,
                                              which_decomp=which_decomp)
                mps[lo] = U.reshape(cl, d1, -1)
                mps[hi] = C.reshape(-1, d2, cr)
            else:         # moving left: U on hi (indices of psi[hi] in theta)
                U, C, info = linalg.factorize(M.T, cutoff=cut, maxdim=mxd, mindim=mnd,
                                              which_decomp=which_decomp)
                mps[hi] = U.T.reshape(-1, d2, cr)
                mps[lo] = C.T.reshape(cl, d1, -1)
            if normalize:
                mps[v] = mps[v] / np.linalg.norm(mps[v])
            envs.site_changed(lo)
            envs.site_changed(hi)
            ortho = v
            res.log.append(RegionLog(sweep, (u, v), energy, info.kept, info.truncerr, info.route,
                                     time.perf_counter() - t_region))
            nreg += 1
            over_time = max_seconds is not None and time.perf_counter() - t0 > max_seconds
            if (max_regions is not None and nreg >= max_regions) or over_time:
                res.energy, res.mps = energy, mps
                res.seconds = time.perf_counter() - t0
                return res
        res.sweep_energies.append(energy)
        res.sweep_maxlinkdim.append(max(a.shape[2] for a in mps))
    res.energy, res.mps = energy, mps
    res.seconds = time.perf_counter() - t0
    return res
