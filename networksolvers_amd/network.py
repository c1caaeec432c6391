"""Device-resident tensor networks on chains: state (MPS), operator (MPO) and the
projected operator with its environment cache.

Restates the parts of ITensorNetworks the sweep needs (``ProjTTN``,
``position``, ``orthogonalize``, ``maxlinkdim``) on a path graph; the layouts
are those of include/nsmi.h (site A[l, s, r], MPO W[wl, wr, s', s],
environment E[bra, mpo, ket]).
"""
from __future__ import annotations

import numpy as np

from . import kernels
from .device import Context, DeviceTensor
from .region_plans import Graph, path_graph


class TTN:
    """Tensor-network state on a path graph, one device tensor per vertex.

    ``ortho_region`` is None while the gauge is unknown (a freshly built state),
    as the reference's TTN whose ortho region is all vertices.  Every write bumps
    a per-vertex stamp that the environment cache checks."""

    def __init__(self, ctx: Context, tensors, graph: Graph | None = None, ortho_region=None):
        self.ctx = ctx
        self.graph = graph or path_graph(len(tensors))
        if not self.graph.is_path():
            raise NotImplementedError("device TTN: only path graphs (chains) are served")
        self.tensors = [t if isinstance(t, DeviceTensor) else ctx.upload(t) for t in tensors]
        self.ortho_region = ortho_region
        self.clock = 0
        self.stamp = [0] * len(self.tensors)

    @classmethod
    def from_arrays(cls, ctx, arrays, graph=None, dtype=None):
        """Upload site tensors; float64 unless any is complex (or ``dtype`` says so)."""
        if dtype is None:
            dtype = np.complex128 if any(np.iscomplexobj(a) for a in arrays) else np.float64
        return cls(ctx, [np.ascontiguousarray(a, dtype=dtype) for a in arrays], graph)

    @property
    def dtype(self):
        return self.tensors[0].dtype

    def __len__(self):
        return len(self.tensors)

    def __getitem__(self, v):
        return self.tensors[v]

    def __setitem__(self, v, t: DeviceTensor):
        self.tensors[v] = t
        self.clock += 1
        self.stamp[v] = self.clock

    def vertices(self):
        return list(self.graph.vertices)

    def linkdims(self):
        return [t.shape[2] for t in self.tensors[:-1]]

    def maxlinkdim(self):
        return max(self.linkdims(), default=1)

    def to_arrays(self):
        return [t.numpy() for t in self.tensors]

    def copy(self):
        c = TTN(self.ctx, [t.copy() for t in self.tensors], self.graph, self.ortho_region)
        return c


class TTNO:
    """Operator (MPO) on a path graph, W[wl, wr, s', s] per vertex, on the device."""

    def __init__(self, ctx: Context, tensors, graph: Graph | None = None):
        self.ctx = ctx
        self.graph = graph or path_graph(len(tensors))
        if not self.graph.is_path():
            raise NotImplementedError("device TTNO: only path graphs (chains) are served")
        self.tensors = [t if isinstance(t, DeviceTensor) else ctx.upload(np.ascontiguousarray(t, dtype=np.float64))
                        for t in tensors]
        self.bond_dims = [t.shape[1] for t in self.tensors[:-1]]
        self._typed = {}

    def tensor(self, v, dtype=np.float64):
        """W of vertex v in the given element type (complex copies cached)."""
        if np.dtype(dtype) == np.float64:
            return self.tensors[v]
        key = (v, np.dtype(dtype).str)
        if key not in self._typed:
            self._typed[key] = self.ctx.upload(self.tensors[v].numpy().astype(dtype))
        return self._typed[key]

    def __len__(self):
        return len(self.tensors)

    def __getitem__(self, v):
        return self.tensors[v]


def orthogonalize(psi: TTN, region) -> TTN:
    """QR walk making the region the orthogonality centre (extracter.jl:6 ->
    ITensorNetworks ``orthogonalize``).  A no-op when the current ortho region
    lies inside the region; otherwise sites left of it become left-orthogonal
    and sites right of it right-orthogonal, pushing R factors inward."""
    lo, hi = min(region), max(region)
    cur = psi.ortho_region
    if cur is not None and all(lo <= v <= hi for v in cur):
        psi.ortho_region = list(region)
        return psi
    N = len(psi)
    left = range(0, lo) if cur is None else range(min(cur), lo)
    right = range(N - 1, hi, -1) if cur is None else range(max(cur), hi, -1)
    for i in left:
        Q, R = kernels.qr(psi[i], 2, False)                  # A[(l s), r] = Q R
        psi[i] = Q
        psi[i + 1] = _absorb_left(R, psi[i + 1])
    for i in right:
        Q, R = kernels.qr(psi[i], 1, True)                   # A[l, (s r)]^T = Q R
        psi[i] = Q
        psi[i - 1] = _absorb_right(psi[i - 1], R)
    psi.ortho_region = list(region)
    return psi


def _absorb_left(R: DeviceTensor, A: DeviceTensor) -> DeviceTensor:
    """R[k, l] . A[l, s, r] -> [k, s, r] via the theta contraction."""
    k, l = R.shape
    Rv = R.copy().reshape_((1, k, l))
    t = kernels.theta(Rv, A)                                  # [1, k, s, r]
    return t.reshape_((k, A.shape[1], A.shape[2]))


def _absorb_right(A: DeviceTensor, Rt: DeviceTensor) -> DeviceTensor:
    """A[l, s, r] . R^T[r, k] -> [l, s, k]."""
    r, k = Rt.shape
    Rv = Rt.copy().reshape_((r, k, 1))
    t = kernels.theta(A, Rv)                                  # [l, s, k, 1]
    return t.reshape_((A.shape[0], A.shape[1], k))


class ProjTTN:
    """Projected operator with cached environments keyed by directed edge (u, v):
    the environment of u's side seen from v (ITensorNetworks ``ProjTTN``).

    ``position(psi, region)`` makes the environments of the edges incident to the
    region current; ``builds`` counts environment contractions (one per region
    step on a sweeping chain)."""

    def __init__(self, H: TTNO, env_precision="fp64"):
        self.H = H
        self.ctx = H.ctx
        self.pos = list(H.graph.vertices)
        self.envs = {}
        self.builds = 0
        self.env_precision = env_precision
        self._one = {}

    def operator(self):
        return self.H

    def sites(self):
        return list(self.pos)

    def incident_edges(self):
        N = len(self.H)
        lo, hi = min(self.pos), max(self.pos)
        out = []
        if lo - 1 >= 0:
            out.append((lo - 1, lo))
        if hi + 1 < N:
            out.append((hi + 1, hi))
        return out

    def _boundary(self, mpo_dim, dtype):
        key = (mpo_dim, np.dtype(dtype).str)
        if key not in self._one:
            self._one[key] = self.ctx.upload(np.ones((1, mpo_dim, 1), dtype=dtype))
        return self._one[key]

    def environment(self, psi: TTN, e):
        """Environment of edge e = (u, v); the trivial one beyond the chain ends."""
        u, v = e
        N = len(self.H)
        if u < 0 or u >= N:
            return self._boundary(1, psi.dtype)
        ent = self.envs.get(e)
        covered = range(0, u + 1) if v == u + 1 else range(u, N)
        if ent is not None:
            t, built = ent
            if all(psi.stamp[x] <= built for x in covered):
                return t
        if v == u + 1:
            prev = self.environment(psi, (u - 1, u))
            t = kernels.env_update(prev, psi[u], self.H[u], 0, self.env_precision)
        else:
            prev = self.environment(psi, (u + 1, u))
            t = kernels.env_update(prev, psi[u], self.H[u], 1, self.env_precision)
        self.builds += 1
        self.envs[e] = (t, psi.clock)
        return t

    def position(self, psi: TTN, region) -> "ProjTTN":
        self.pos = sorted(region)
        lo, hi = self.pos[0], self.pos[-1]
        for e in list(self.envs):               # environments of internal edges are stale
            if lo <= e[0] <= hi and lo <= e[1] <= hi:
                del self.envs[e]
        for e in self.incident_edges():
            self.environment(psi, e)
        return self

    def left_right(self, psi: TTN):
        lo, hi = min(self.pos), max(self.pos)
        return self.environment(psi, (lo - 1, lo)), self.environment(psi, (hi + 1, hi))


def expect_mpo(psi: TTN, H: TTNO, precision="fp64") -> float:
    """<psi|H|psi> / <psi|psi> by a full left-to-right environment pass in FP64
    (on a tree: the Rayleigh quotient at an orthogonality centre)."""
    if not isinstance(psi, TTN):
        from . import qn, tree
        if isinstance(psi, qn.QNMPS):
            return qn.expect_mpo(psi, H)
        return tree.expect_mpo(psi, H)
    ctx = H.ctx
    E = ctx.upload(np.ones((1, 1, 1), dtype=psi.dtype))
    for i in range(len(psi)):
        E = kernels.env_update(E, psi[i], H[i], 0, precision)
    val = E.numpy().reshape(-1)[0]
    return float(np.real(val)) / norm2(psi)


def expect_local(psi, op) -> dict:
    """<psi|op_v|psi> / <psi|psi> for every vertex v of a chain or tree state
    (ITensorNetworks ``expect`` with a one-site operator): the orthogonality
    centre of a copy walks site by site and the value is read at the centre."""
    if not isinstance(psi, TTN):
        from . import tree
        return tree.expect_local(psi, op)
    c = psi.copy()
    opd = psi.ctx.upload(np.ascontiguousarray(np.asarray(op), dtype=psi.dtype))
    out = {}
    for v in range(len(c)):
        orthogonalize(c, [v])
        A = c[v]
        B = kernels.tensordot(A, opd, [1], [1])                 # [l, r, s']
        B = kernels.permute(B, [0, 2, 1])                       # [l, s', r]
        out[v] = float(np.real(A.inner(B) / A.inner(A)))
    return out


def norm2(psi: TTN) -> float:
    """<psi|psi> by a transfer-operator pass (identity MPO)."""
    ctx = psi.ctx
    E = ctx.upload(np.ones((1, 1, 1), dtype=psi.dtype))
    for i in range(len(psi)):
        d = psi[i].shape[1]
        E = kernels.env_update(E, psi[i], ctx.upload(np.eye(d).reshape(1, 1, d, d)), 0)
    return float(np.real(E.numpy().reshape(-1)[0]))


def inner(a: TTN, b: TTN):
    """<a|b> (host contraction of downloaded tensors; for checks and observables)."""
    E = np.ones((1, 1), dtype=np.complex128)
    for x, y in zip(a.to_arrays(), b.to_arrays()):
        E = np.tensordot(x.conj(), np.tensordot(E, y, axes=([1], [0])), axes=([0, 1], [0, 1]))
    v = E[0, 0]
    return v if (np.iscomplexobj(a.tensors[0].numpy()) or np.iscomplexobj(b.tensors[0].numpy())) else v.real
