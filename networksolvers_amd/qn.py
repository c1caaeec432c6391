"""U(1) quantum-number (block-sparse) chains on the device.

The reference runs the same sweeps on QN-conserving ITensors (siteinds with
conserve_qns=true; README.md QN timings, examples/dmrg.jl conserve_qns): every
tensor is a set of dense blocks, one per charge-consistent combination of index
sectors, and every contraction is a sum over compatible block pairs
(operator_map.jl:3-10 "per compatible block pair").  Here the host works out the
block bookkeeping and the device executes it as block programs
(``nsmi_bsop_create``): grouped GEMM launches in which each output block sums its
segments in one pass, grouped linear combinations for the MPO mixing, grouped
copies for sector assembly.  The truncated split is ``nsmi_factorize_sectors``.

Conventions.  Charges are integers (2 Sz for spin states).  A site tensor
A[l, s, r] has arrows (in, in, out): q_l + q_s = q_r; the two-site tensor
theta[l, s1, s2, r]: q_l + q_s1 + q_s2 = q_r; an environment E[bra, w, ket]:
q_bra - q_ket = c_w, where c_w is the MPO link charge, propagated from the left
boundary by c_wr = c_wl + q_s' - q_s over the MPO's nonzeros.  Site and MPO
indices have one sector per basis state, so every block is a (link x link)
matrix, stored row-major, blocks concatenated in one flat FP64 buffer.
"""
from __future__ import annotations

import ctypes as C
import itertools
import math

import numpy as np

from . import _lib, kernels
from .device import Context, DeviceTensor
from .region_plans import path_graph


# ------------------------------------------------------------------ indices
class QNIndex:
    """Sectors (charge, dim) in a fixed order; dense offsets follow that order."""

    def __init__(self, sectors):
        self.q = [int(a) for a, n in sectors]
        self.n = [int(n) for a, n in sectors]
        if any(n <= 0 for n in self.n):
            raise ValueError("sector dimensions must be positive")
        self.off = [int(x) for x in np.concatenate([[0], np.cumsum(self.n)])[:-1]] if self.n else []
        self.dim = int(sum(self.n))
        self._by_q = {}
        for i, q in enumerate(self.q):
            self._by_q.setdefault(q, []).append(i)

    def find(self, q):
        """The sector with charge q (link indices have unique charges), or None."""
        ids = self._by_q.get(int(q))
        return ids[0] if ids else None

    def __len__(self):
        return len(self.q)

    def __repr__(self):
        return f"QNIndex({list(zip(self.q, self.n))})"


def spin_site_index(d):
    """Spin S = (d-1)/2 basis m = S, S-1, ..., -S ("Up" first) with charge 2m."""
    return QNIndex([(int(round(2 * ((d - 1) / 2 - i))), 1) for i in range(d)])


# ------------------------------------------------------------------ tensors
_KEYS_CACHE = {}


class QNTensor:
    """A block-sparse tensor: ``keys`` (tuples of sector ids) in storage order, each
    block dense with the sector dims, all in one flat device buffer."""

    def __init__(self, ctx: Context, indices, dirs, flux=0, keys=None, data: DeviceTensor | None = None,
                 zero=True):
        self.ctx = ctx
        self.indices = list(indices)
        self.dirs = list(dirs)
        self.flux = int(flux)
        if keys is None:
            keys = self._allowed_keys()
        self.keys = [tuple(k) for k in keys]
        self.pos = {k: i for i, k in enumerate(self.keys)}
        self.shapes = [tuple(ix.n[i] for ix, i in zip(self.indices, k)) for k in self.keys]
        sizes = [math.prod(s) for s in self.shapes]
        self.offs = [0] + list(itertools.accumulate(sizes))[:-1] if sizes else []
        self.size = int(sum(sizes))
        if data is None:   # zeros, or an unset buffer for an output every block of which gets written
            data = ctx.upload(np.zeros(max(self.size, 1))) if zero else ctx.empty((max(self.size, 1),))
        self.data = data

    def _allowed_keys(self):
        """Every charge-allowed key in lexicographic order: the leading indices run
        over all their sectors and the last one takes the sectors of the charge
        that balances the flux (cached per charge structure: the keys depend only
        on the sector charges, the arrows and the flux)."""
        ck = (tuple(tuple(ix.q) for ix in self.indices), tuple(self.dirs), self.flux)
        keys = _KEYS_CACHE.get(ck)
        if keys is None:
            keys = self._enumerate_keys()
            if len(_KEYS_CACHE) > 4096:
                _KEYS_CACHE.clear()
            _KEYS_CACHE[ck] = keys
        return keys

    def _enumerate_keys(self):
        *lead, last = self.indices
        *dl, d_last = self.dirs
        keys = []
        for k in itertools.product(*[range(len(ix)) for ix in lead]):
            q = self.flux - sum(d * ix.q[i] for d, ix, i in zip(dl, lead, k))
            if q % d_last:
                continue
            for j in last._by_q.get(q // d_last, ()):
                keys.append(k + (j,))
        return keys

    def rows_cols(self, key):
        s = self.shapes[self.pos[key]]
        return s[0], math.prod(s[1:])

    def block_off(self, key):
        return self.offs[self.pos[key]]

    def has(self, key):
        return key in self.pos

    @classmethod
    def from_dense(cls, ctx, arr, indices, dirs, flux=0, atol=0.0):
        t = cls(ctx, indices, dirs, flux)
        flat = np.zeros(max(t.size, 1))
        covered = np.zeros(arr.shape, dtype=bool)
        for k, off, shp in zip(t.keys, t.offs, t.shapes):
            sl = tuple(slice(ix.off[i], ix.off[i] + ix.n[i]) for ix, i in zip(t.indices, k))
            flat[off:off + int(np.prod(shp))] = np.asarray(arr[sl], dtype=np.float64).reshape(-1)
            covered[sl] = True
        if np.any(np.abs(arr[~covered]) > atol):
            raise ValueError("dense array has weight outside the charge-allowed blocks")
        t.data = ctx.upload(flat)
        return t

    def to_dense(self):
        flat = self.data.numpy()
        out = np.zeros([ix.dim for ix in self.indices])
        for k, off, shp in zip(self.keys, self.offs, self.shapes):
            sl = tuple(slice(ix.off[i], ix.off[i] + ix.n[i]) for ix, i in zip(self.indices, k))
            out[sl] = flat[off:off + int(np.prod(shp))].reshape(shp)
        return out

    def copy(self):
        return QNTensor(self.ctx, self.indices, self.dirs, self.flux, self.keys, self.data.copy())

    def norm(self):
        return self.data.norm()

    @property
    def h(self):
        """The flat buffer's handle: Krylov solvers act on the block vector directly."""
        return self.data.h

    @property
    def shape(self):
        return (self.size,)

    @property
    def dtype(self):
        return np.dtype(np.float64)

    def scale_(self, a):
        self.data.scale_(a)
        return self


# ------------------------------------------------------------ block programs
class _Program:
    """Collects the steps of one block program (buffer ids: 0 = x, 1 = y, 2.. =
    operands, then scratch)."""

    def __init__(self):
        self.operands = []
        self.scratch = []
        self.steps = []
        self.gitems, self.segs, self.alpha = [], [], []
        self.litems, self.terms, self.coeff = [], [], []
        self.citems = []

    def operand(self, t: DeviceTensor):
        for i, o in enumerate(self.operands):
            if o is t:
                return 2 + i
        self.operands.append(t)
        return 1 + len(self.operands)

    def scratch_buf(self, n):
        self.scratch.append(int(n))
        return None  # resolved in build (ids after the operands)

    # The item tables are kept flat (7 ints per GEMM item and segment, 5 per linear
    # combination item, 2 per term, 8 per copy) so that build converts them at once.
    def gemm(self, ta, tb, items):
        """items: [(c_buf, c_off, ldc, M, N, [(a_buf, a_off, lda, b_buf, b_off, ldb, K, alpha), ...])]"""
        gi, sgl, al = self.gitems, self.segs, self.alpha
        first = len(gi) // 7
        nseg = len(al)
        for c_buf, c_off, ldc, M, N, sg in items:
            gi += (c_buf, c_off, ldc, M, N, nseg, len(sg))
            for seg in sg:
                sgl += seg[:7]
                al.append(seg[7])
            nseg += len(sg)
        self.steps.append((0, int(ta), int(tb), first, len(items), 0))

    def lincomb(self, items):
        """items: [(out_buf, out_off, n, [(src_buf, src_off, coeff), ...])]"""
        li, te, co = self.litems, self.terms, self.coeff
        first = len(li) // 5
        nt = len(co)
        for ob, oo, n, tl in items:
            li += (ob, oo, n, nt, len(tl))
            for sb, so, c in tl:
                te += (sb, so)
                co.append(c)
            nt += len(tl)
        self.steps.append((1, 0, 0, first, len(items), 0))

    def copy(self, items):
        """items: [(src_buf, src_off, src_ld, dst_buf, dst_off, dst_ld, rows, cols)]"""
        first = len(self.citems) // 8
        for it in items:
            self.citems += it
        self.steps.append((2, 0, 0, first, len(items), 0))

    def sid(self, j):
        """Buffer id of scratch buffer j (valid once all operands are registered)."""
        return 2 + len(self.operands) + j

    def build(self, ctx: Context, x_elems, y_elems, precision="fp64"):
        keep = []

        def arr(rows, w, dt=np.int64, ct=C.c_int64):   # rows of w entries as one contiguous array
            a = np.asarray(rows, dtype=dt).reshape(-1) if len(rows) else np.zeros(w, dtype=dt)
            keep.append(a)
            return a.ctypes.data_as(C.POINTER(ct))

        ops = (C.c_void_p * max(len(self.operands), 1))(*[o.h.value for o in self.operands])
        scr = (C.c_int64 * max(len(self.scratch), 1))(*self.scratch)
        h = C.c_void_p()
        ctx.check(ctx.lib.nsmi_bsop_create(
            ctx.h, len(self.operands), ops, len(self.scratch), scr, int(x_elems), int(y_elems), len(self.steps),
            arr(self.steps, 6), arr(self.gitems, 7), arr(self.segs, 7), arr(self.alpha, 1, np.float64, C.c_double),
            arr(self.litems, 5), arr(self.terms, 2), arr(self.coeff, 1, np.float64, C.c_double),
            arr(self.citems, 8), kernels._prec(precision), C.byref(h)))
        return BlockProgram(ctx, h, self.operands, precision)


class BlockProgram(kernels._OperatorBase):
    """A built block program; an effective Hamiltonian when x and y have one size.
    ``precision`` "fp32" runs its GEMMs, mixing and copies in FP32 (FP32 MFMA) on
    FP32 copies of the operands, with x and y converted per run."""

    def __init__(self, ctx, h, keep, precision="fp64"):
        self.ctx = ctx
        self.h = h
        self._keep = tuple(keep)
        self.precision = precision

    def run(self, x: DeviceTensor | None, y: DeviceTensor | None):
        self.ctx.check(self.ctx.lib.nsmi_bsop_run(self.ctx.h, self.h, x.h if x is not None else None,
                                                  y.h if y is not None else None))


def _run_once(prog: _Program, ctx, x, y):
    op = prog.build(ctx, x.size if x is not None else 0, y.size if y is not None else 0)
    try:
        op.run(x.data if x is not None else None, y.data if y is not None else None)
    finally:
        op.close()


# ---------------------------------------------------------------- operators
class QNMPO:
    """Chain MPO W[wl, wr, s', s] (host, FP64) with its MPO link charges inferred
    by propagation from the left boundary."""

    def __init__(self, ctx: Context, W, site: QNIndex):
        self.ctx = ctx
        self.W = [np.asarray(w, dtype=np.float64) for w in W]
        self.site = site
        qs = site.q
        N = len(self.W)
        self.c = [np.zeros(self.W[0].shape[0], dtype=np.int64)]
        for i, w in enumerate(self.W):
            cl = self.c[-1]
            cr = np.full(w.shape[1], np.iinfo(np.int64).min, dtype=np.int64)
            for a, b, sp, s in zip(*np.nonzero(w)):
                v = cl[a] + qs[sp] - qs[s]
                if cr[b] == np.iinfo(np.int64).min:
                    cr[b] = v
                elif cr[b] != v:
                    raise ValueError(f"MPO site {i} does not conserve the charge")
            cr[cr == np.iinfo(np.int64).min] = 0
            self.c.append(cr)
        if N and any(self.c[-1] != 0):
            raise ValueError("MPO right boundary must carry charge 0")
        self.links = [QNIndex([(int(q), 1) for q in c]) for c in self.c]

    def __len__(self):
        return len(self.W)


def qn_product_mps(ctx, states, site: QNIndex):
    """Product state with site i in basis state states[i]: link charges accumulate."""
    links = [QNIndex([(0, 1)])]
    q = 0
    for s in states:
        q += site.q[s]
        links.append(QNIndex([(q, 1)]))
    tensors = []
    for i, s in enumerate(states):
        A = QNTensor(ctx, [links[i], site, links[i + 1]], [1, 1, -1], 0, keys=[(0, s, 0)])
        A.data = ctx.upload(np.ones(1))
        tensors.append(A)
    return QNMPS(ctx, tensors, links, site)


class QNMPS:
    """QN chain state; every site tensor is an isometry for a product state, so the
    state starts with a known (trivial) gauge."""

    def __init__(self, ctx, tensors, links, site, ortho_region=None):
        self.ctx = ctx
        self.tensors = list(tensors)
        self.links = list(links)
        self.site = site
        self.graph = path_graph(len(tensors))
        self.ortho_region = ortho_region if ortho_region is not None else [len(tensors) - 1]
        self.clock = 0
        self.stamp = [0] * len(tensors)

    def __len__(self):
        return len(self.tensors)

    def __getitem__(self, i):
        return self.tensors[i]

    def __setitem__(self, i, t):
        self.tensors[i] = t
        self.clock += 1
        self.stamp[i] = self.clock

    @property
    def dtype(self):
        return np.dtype(np.float64)

    def linkdims(self):
        return [self.links[i + 1].dim for i in range(len(self) - 1)]

    def maxlinkdim(self):
        return max(self.linkdims(), default=1)

    def to_arrays(self):
        return [t.to_dense() for t in self.tensors]

    def copy(self):
        c = QNMPS(self.ctx, [t.copy() for t in self.tensors], self.links, self.site, list(self.ortho_region))
        return c


def _boundary_env(ctx, link: QNIndex, mpo_link: QNIndex):
    E = QNTensor(ctx, [link, mpo_link, link], [1, -1, -1], 0)
    E.data = ctx.upload(np.ones(max(E.size, 1)))
    return E


class _Mix:
    """The first two steps of an environment update, planned on the host: X = E.A
    (left) or A.E^T (right) as one grouped GEMM, then the MPO mix Y = sum_c c X as
    one grouped linear combination.  ``Y`` maps a block key to its offset in the
    Y scratch buffer and ``ysize`` to its element count."""

    def __init__(self, left, xs, X, X_off, ys, ysize):
        self.left, self.xs, self.X, self.X_off, self.ys, self.ysize = left, xs, X, X_off, ys, ysize
        self.Y, off = {}, 0
        for k in ys:
            self.Y[k] = off
            off += ysize[k]
        self.Y_off = off

    def emit(self, p, xE, xA):
        """Appends the two steps to program p (every operand already registered);
        returns the buffer id of Y."""
        j0 = len(p.scratch)
        p.scratch_buf(self.X_off)
        p.scratch_buf(self.Y_off)
        sX, sY = p.sid(j0), p.sid(j0 + 1)
        if self.left:
            p.gemm(False, False, [(sX, self.X[k], nb, na_p, nb, [(xE, eo, na, xA, ao, nb, na, 1.0)])
                                  for (k, eo, ao, na_p, na, nb) in self.xs])
        else:
            p.gemm(False, True, [(sX, self.X[k], nbp, na, nbp, [(xA, ao, nb, xE, eo, nb, nb, 1.0)])
                                 for (k, ao, eo, na, nb, nbp) in self.xs])
        p.lincomb([(sY, self.Y[k], self.ysize[k], [(sX, self.X[xk], c) for xk, c in tl])
                   for k, tl in self.ys.items()])
        return sY


def _w_nonzeros(W, out_axis):
    """Nonzero entries of the MPO tensor W[wl, wr, s', s] grouped by (the other MPO
    index, s): lists of (out index, s', c) in the order of the loops over the output
    MPO index, then s'."""
    nz = {}
    Dl, Dr, dp, d = W.shape
    for i in range(Dl if out_axis == 1 else Dr):
        for s in range(d):
            lst = []
            for o in range(Dr if out_axis == 1 else Dl):
                for sp in range(dp):
                    c = float(W[i, o, sp, s] if out_axis == 1 else W[o, i, sp, s])
                    if c != 0.0:
                        lst.append((o, sp, c))
            nz[(i, s)] = lst
    return nz


def _mix_left(E: QNTensor, A: QNTensor, W, cl) -> _Mix:
    """Y[(ap, sp, wp, b)] = sum_{w, s} W[w, wp, sp, s] (E[ap, w, a] A[a, s, b]), blocks
    l.n[ap] x r.n[b], for A[l, s, r] and the environment E[l', w, l] on its left."""
    l, site, r = A.indices
    # per left sector a: the environment blocks E[a', w, a] that exist, in w order
    left = []
    for a in range(len(l)):
        lst = []
        for w, cw in enumerate(cl):
            ap = l.find(l.q[a] + cw)
            i = None if ap is None else E.pos.get((ap, w, a))
            if i is not None:
                lst.append((w, ap, E.offs[i]))
        left.append(lst)
    X, X_off, xs = {}, 0, []
    for (a, s, b), ao in zip(A.keys, A.offs):
        for w, ap, eo in left[a]:
            X[(ap, w, s, b)] = X_off
            xs.append(((ap, w, s, b), eo, ao, l.n[ap], l.n[a], r.n[b]))
            X_off += l.n[ap] * r.n[b]
    wnz = _w_nonzeros(W, 1)
    ys = {}
    for k in X:
        ap, w, s, b = k
        for wp, sp, c in wnz[(w, s)]:
            kk = (ap, sp, wp, b)
            lst = ys.get(kk)
            if lst is None:
                ys[kk] = [(k, c)]
            else:
                lst.append((k, c))
    return _Mix(True, xs, X, X_off, ys, {k: l.n[k[0]] * r.n[k[3]] for k in ys})


def _mix_right(E: QNTensor, A: QNTensor, W, cr) -> _Mix:
    """Y[(a, w, sp, bp)] = sum_{v, s} W[w, v, sp, s] (A[a, s, b] E[bp, v, b]), blocks
    l.n[a] x r.n[bp], for A[l, s, r] and the environment E[r', v, r] on its right."""
    l, site, r = A.indices
    # per right sector b: the environment blocks E[b', v, b] that exist, in v order
    right = []
    for b in range(len(r)):
        lst = []
        for v, cv in enumerate(cr):
            bp = r.find(r.q[b] + cv)
            i = None if bp is None else E.pos.get((bp, v, b))
            if i is not None:
                lst.append((v, bp, E.offs[i]))
        right.append(lst)
    X, X_off, xs = {}, 0, []
    for (a, s, b), ao in zip(A.keys, A.offs):
        for v, bp, eo in right[b]:
            X[(a, s, v, bp)] = X_off
            xs.append(((a, s, v, bp), ao, eo, l.n[a], r.n[b], r.n[bp]))
            X_off += l.n[a] * r.n[bp]
    wnz = _w_nonzeros(W, 0)
    ys = {}
    for k in X:
        a, s, v, bp = k
        for w, sp, c in wnz[(v, s)]:
            kk = (a, w, sp, bp)
            lst = ys.get(kk)
            if lst is None:
                ys[kk] = [(k, c)]
            else:
                lst.append((k, c))
    return _Mix(False, xs, X, X_off, ys, {k: l.n[k[0]] * r.n[k[3]] for k in ys})


def env_left(E: QNTensor, A: QNTensor, W, cl, cr, ctx) -> QNTensor:
    """E'[b', w', b] = sum conj(A[a', s', b']) E[a', w, a] W[w, w', s', s] A[a, s, b]
    (extracter.jl:14 -> ProjTTN environment) as a block program: X = E.A (GEMM),
    Y = W-mix (LINCOMB), E' = A^T.Y (GEMM with segments over (a', s'))."""
    l, site, r = A.indices
    mw_r = QNIndex([(int(q), 1) for q in cr])
    out = QNTensor(ctx, [r, mw_r, r], [1, -1, -1], 0, zero=False)
    mix = _mix_left(E, A, W, cl)
    p = _Program()
    xE, xA = p.operand(E.data), 0
    sY = mix.emit(p, xE, xA)
    Y = mix.Y
    by_last = {}   # A's blocks grouped by their right sector, in A's key order
    for (ap, sp, bp), ao in zip(A.keys, A.offs):
        by_last.setdefault(bp, []).append((ap, sp, ao))
    items = []
    for (bp, wp, b), oo in zip(out.keys, out.offs):
        sg = []
        for ap, sp, ao in by_last.get(bp, ()):
            y = Y.get((ap, sp, wp, b))
            if y is not None:
                sg.append((xA, ao, r.n[bp], sY, y, r.n[b], l.n[ap], 1.0))
        items.append((1, oo, r.n[b], r.n[bp], r.n[b], sg))
    p.gemm(True, False, items)
    _run_once(p, ctx, A, out)
    return out


def env_right(E: QNTensor, A: QNTensor, W, cl, cr, ctx) -> QNTensor:
    """E'[a', w, a] = sum conj(A[a', s', b']) E[b', v, b] W[w, v, s', s] A[a, s, b]:
    X = A.E^T (GEMM), Y = W-mix (LINCOMB), E' = A.Y^T (GEMM, segments over (s', b'))."""
    l, site, r = A.indices
    mw_l = QNIndex([(int(q), 1) for q in cl])
    out = QNTensor(ctx, [l, mw_l, l], [1, -1, -1], 0, zero=False)
    mix = _mix_right(E, A, W, cr)
    p = _Program()
    xE, xA = p.operand(E.data), 0
    sY = mix.emit(p, xE, xA)
    Y = mix.Y
    by_first = {}   # A's blocks grouped by their left sector, in A's key order
    for (app, sp, bp), ao in zip(A.keys, A.offs):
        by_first.setdefault(app, []).append((sp, bp, ao))
    items = []
    for (ap, w, a), oo in zip(out.keys, out.offs):
        sg = []
        for sp, bp, ao in by_first.get(ap, ()):
            y = Y.get((a, w, sp, bp))
            if y is not None:
                sg.append((xA, ao, r.n[bp], sY, y, r.n[bp], r.n[bp], 1.0))
        items.append((1, oo, l.n[a], l.n[ap], l.n[a], sg))
    p.gemm(False, True, items)
    _run_once(p, ctx, A, out)
    return out


def theta(A: QNTensor, B: QNTensor, ctx) -> QNTensor:
    """theta[l, s1, s2, r] = sum_m A[l, s1, m] B[m, s2, r] (extracter.jl:7) over all
    charge-allowed blocks of the outer links (a block without a contributing pair
    is zero)."""
    l, s1i, m = A.indices
    _, s2i, r = B.indices
    th = QNTensor(ctx, [l, s1i, s2i, r], [1, 1, 1, -1], 0, zero=False)
    p = _Program()
    xB = p.operand(B.data)
    items = []
    for (a, s1, s2, b) in th.keys:
        mm = m.find(l.q[a] + s1i.q[s1])
        sg = []
        if mm is not None and A.has((a, s1, mm)) and B.has((mm, s2, b)):
            sg.append((0, A.block_off((a, s1, mm)), m.n[mm], xB, B.block_off((mm, s2, b)), r.n[b], m.n[mm], 1.0))
        items.append((1, th.block_off((a, s1, s2, b)), r.n[b], l.n[a], r.n[b], sg))
    p.gemm(False, False, items)
    _run_once(p, ctx, A, th)
    return th


_W12_NZ = {}


def _w12_nonzeros(W1, W2):
    """Nonzero entries of W12[w, s1, s2, s1', s2', v] = sum_m W1[w, m, s1', s1] W2[m, v, s2', s2],
    grouped by (w, s1, s2) in row-major order; cached per pair of MPO tensors."""
    key = (W1.shape, W2.shape, W1.tobytes(), W2.tobytes())
    nz = _W12_NZ.get(key)
    if nz is None:
        W12 = np.einsum("wmas,mvbt->wstabv", W1, W2)   # [w, s1, s2, s1', s2', v]
        nz = {}
        for w, s1, s2, a1, a2, v in zip(*np.nonzero(W12)):
            nz.setdefault((int(w), int(s1), int(s2)), []).append((int(a1), int(a2), int(v),
                                                                  float(W12[w, s1, s2, a1, a2, v])))
        if len(_W12_NZ) > 256:
            _W12_NZ.clear()
        _W12_NZ[key] = nz
    return nz


def heff2(L: QNTensor, W1, W2, R: QNTensor, th: QNTensor, cL, cR, ctx, precision="fp64") -> BlockProgram:
    """Two-site effective Hamiltonian (operator_map.jl:3-10) on theta's block space:
    T1 = L.theta (GEMM), T3 = (W1 W2)-mix (LINCOMB), y = T3.R^T (GEMM, segments over
    the right MPO link); in FP64 or the FP32 form."""
    l, s1i, s2i, r = th.indices
    nz = _w12_nonzeros(W1, W2)
    p = _Program()
    xL, xR = p.operand(L.data), p.operand(R.data)
    ln, rn = l.n, r.n
    # per left sector a: the environment blocks L[a', w, a] that exist, in w order
    left = []
    for a in range(len(l)):
        lst = []
        for w, cw in enumerate(cL):
            ap = l.find(l.q[a] + cw)
            i = None if ap is None else L.pos.get((ap, w, a))
            if i is not None:
                lst.append((w, ap, L.offs[i]))
        left.append(lst)
    T1, T1_off, t1 = {}, 0, []
    for (a, s1, s2, b), to in zip(th.keys, th.offs):
        nb = rn[b]
        for w, ap, lo in left[a]:
            k = (ap, w, s1, s2, b)
            T1[k] = T1_off
            t1.append((k, lo, to, ln[ap], ln[a], nb))
            T1_off += ln[ap] * nb
    t3 = {}
    for k in T1:
        ap, w, s1, s2, b = k
        for a1, a2, v, c in nz.get((w, s1, s2), ()):
            kk = (ap, a1, a2, v, b)
            lst = t3.get(kk)
            if lst is None:
                t3[kk] = [(k, c)]
            else:
                lst.append((k, c))
    T3, T3_off = {}, 0
    for k in t3:
        T3[k] = T3_off
        T3_off += ln[k[0]] * rn[k[4]]
    p.scratch_buf(T1_off)
    p.scratch_buf(T3_off)
    s1b, s3b = p.sid(0), p.sid(1)
    p.gemm(False, False, [(s1b, T1[k], nb, nap, nb, [(xL, lo, na, 0, to, nb, na, 1.0)])
                          for (k, lo, to, nap, na, nb) in t1])
    p.lincomb([(s3b, T3[k], ln[k[0]] * rn[k[4]], [(s1b, T1[tk], c) for tk, c in tl]) for k, tl in t3.items()])
    # per right sector b': the environment blocks R[b', v, b] that exist, in v order
    right = []
    for bp in range(len(r)):
        lst = []
        for v, cv in enumerate(cR):
            b = r.find(r.q[bp] - cv)
            i = None if b is None else R.pos.get((bp, v, b))
            if i is not None:
                lst.append((v, b, R.offs[i]))
        right.append(lst)
    items = []
    for (ap, a1, a2, bp), to in zip(th.keys, th.offs):
        sg = []
        for v, b, ro in right[bp]:
            t = T3.get((ap, a1, a2, v, b))
            if t is not None:
                sg.append((s3b, t, rn[b], xR, ro, rn[b], rn[b], 1.0))
        items.append((1, to, rn[bp], ln[ap], rn[bp], sg))
    p.gemm(False, True, items)
    return p.build(ctx, th.size, th.size, precision)


def heff1(L: QNTensor, W, R: QNTensor, A: QNTensor, cL, cR, ctx, precision="fp64") -> BlockProgram:
    """One-site effective Hamiltonian on A's block space (the 1-site regions of
    operator_map.jl:3-10): the left environment update's first two steps give
    Y[(a', s', v, b)] = sum W[w, v, s', s] L[a', w, a] x[a, s, b], then
    y[a', s', b'] = sum_{v, b} Y R[b', v, b]^T (GEMM with segments over the right
    MPO link)."""
    l, site, r = A.indices
    mix = _mix_left(L, A, W, cL)
    p = _Program()
    xL, xR = p.operand(L.data), p.operand(R.data)
    sY = mix.emit(p, xL, 0)
    items = []
    for (ap, sp, bp) in A.keys:
        sg = []
        for v in range(len(cR)):
            b = r.find(r.q[bp] - cR[v])
            if b is None or (ap, sp, v, b) not in mix.Y or not R.has((bp, v, b)):
                continue
            sg.append((sY, mix.Y[(ap, sp, v, b)], r.n[b], xR, R.block_off((bp, v, b)), r.n[b], r.n[b], 1.0))
        items.append((1, A.block_off((ap, sp, bp)), r.n[bp], l.n[ap], r.n[bp], sg))
    p.gemm(False, True, items)
    return p.build(ctx, A.size, A.size, precision)


def factorize(th: QNTensor, rows_trailing: bool, *, cutoff=0.0, maxdim=None, mindim=1, which_decomp=None, ctx=None):
    """Truncated split of theta[l, s1, s2, r] over the new link between s1 and s2
    (inserter.jl:23 with a QN ITensor): sector q = q_l + q_s1 gathers rows (l, s1)
    and columns (s2, r); ``nsmi_factorize_sectors`` decomposes, truncates globally
    and keeps per sector; returns (left tensor [l, s1, m], right tensor [m, s2, r],
    new link, TruncInfo).  rows_trailing puts the isometry on the right tensor."""
    ctx = ctx or th.ctx
    l, s1i, s2i, r = th.indices
    rows, cols = {}, {}
    for (a, s1, s2, b) in th.keys:
        q = l.q[a] + s1i.q[s1]
        rows.setdefault(q, {})[(a, s1)] = None
        cols.setdefault(q, {})[(s2, b)] = None
    sectors = sorted(q for q in rows if q in cols)
    ro, co, R_, C_, M_off = {}, {}, [], [], []
    tot = 0
    for q in sectors:
        off = 0
        for (a, s1) in sorted(rows[q]):
            ro[(a, s1)] = off
            off += l.n[a]
        R_.append(off)
        off = 0
        for (s2, b) in sorted(cols[q]):
            co[(s2, b)] = off
            off += r.n[b]
        C_.append(off)
        M_off.append(tot)
        tot += R_[-1] * C_[-1]
    qi = {q: i for i, q in enumerate(sectors)}
    Mall = ctx.empty((max(tot, 1),))   # every (row, column) block of a sector is a block of theta
    p = _Program()
    yM = p.operand(Mall)
    items = []
    for (a, s1, s2, b) in th.keys:
        i = qi[l.q[a] + s1i.q[s1]]
        items.append((0, th.block_off((a, s1, s2, b)), r.n[b], yM, M_off[i] + ro[(a, s1)] * C_[i] + co[(s2, b)], C_[i],
                      l.n[a], r.n[b]))
    p.copy(items)
    _run_once(p, ctx, th, None)
    Uall, Call, ks, ti = _factorize_sectors(ctx, Mall, M_off, R_, C_, rows_trailing, cutoff, maxdim, mindim,
                                            which_decomp)
    mid = QNIndex([(q, k) for q, k in zip(sectors, ks) if k > 0])
    left = QNTensor(ctx, [l, s1i, mid], [1, 1, -1], 0, zero=False)
    right = QNTensor(ctx, [mid, s2i, r], [1, 1, -1], 0, zero=False)
    # offsets of U_q / C_q in the concatenated outputs
    uoff, coff, uo, cof = {}, {}, 0, 0
    for i, q in enumerate(sectors):
        uoff[q], coff[q] = uo, cof
        k = ks[i]
        uo += k * (C_[i] if rows_trailing else R_[i])
        cof += k * (R_[i] if rows_trailing else C_[i])
    p = _Program()
    xU, xC = p.operand(Uall), p.operand(Call)
    yl, yr = p.operand(left.data), p.operand(right.data)
    items = []
    for (a, s1, m) in left.keys:
        q = mid.q[m]
        i, k = qi[q], mid.n[m]
        src = xC if rows_trailing else xU
        base = coff[q] if rows_trailing else uoff[q]
        items.append((src, base + ro[(a, s1)] * k, k, yl, left.block_off((a, s1, m)), k, l.n[a], k))
    for (m, s2, b) in right.keys:
        q = mid.q[m]
        i, k = qi[q], mid.n[m]
        src = xU if rows_trailing else xC
        base = uoff[q] if rows_trailing else coff[q]
        items.append((src, base + co[(s2, b)], C_[i], yr, right.block_off((m, s2, b)), r.n[b], k, r.n[b]))
    p.copy(items)
    op = p.build(ctx, 0, 0)
    try:
        op.run(None, None)
    finally:
        op.close()
    return left, right, mid, ti


def _factorize_sectors(ctx, Mall, offsets, rows, cols, rows_trailing, cutoff, maxdim, mindim, which_decomp):
    """nsmi_factorize_sectors: (U_all, C_all, kept per sector, TruncInfo)."""
    n = len(offsets)
    route = _lib.ROUTE["auto" if which_decomp is None else which_decomp]
    t = _lib.Trunc(float(cutoff), int(maxdim) if maxdim is not None else 0, int(mindim), route)
    kept = (C.c_int64 * max(n, 1))()
    info = _lib.TruncInfo()
    Uh, Ch = C.c_void_p(), C.c_void_p()
    arr = lambda v: (C.c_int64 * max(len(v), 1))(*v)    # noqa: E731
    ctx.check(ctx.lib.nsmi_factorize_sectors(ctx.h, Mall.h, n, arr(offsets), arr(rows), arr(cols),
                                             int(bool(rows_trailing)), C.byref(t), C.byref(Uh), C.byref(Ch), kept,
                                             C.byref(info)))
    ti = kernels.TruncInfo(info.kept, info.full, info.truncerr, info.total, _lib.ROUTE_NAME.get(info.route, "?"))
    return DeviceTensor(ctx, Uh), DeviceTensor(ctx, Ch), [int(kept[i]) for i in range(n)], ti


# ------------------------------------------------------------ projected operator
class QNProjTTN:
    """Environments of a QN chain per directed edge (ProjTTN restricted to a chain),
    rebuilt when a site on their side changed."""

    def __init__(self, H: QNMPO, ctx: Context):
        self.H = H
        self.ctx = ctx
        self.pos = list(range(len(H)))
        self.envs = {}
        self.builds = 0
        self.env_precision = "fp64"

    def operator(self):
        return self.H

    def sites(self):
        return list(self.pos)

    def environment(self, psi: QNMPS, e):
        u, v = e
        N = len(self.H)
        if u < 0:
            return _boundary_env(self.ctx, psi.links[0], self.H.links[0])
        if u >= N:
            return _boundary_env(self.ctx, psi.links[N], self.H.links[N])
        ent = self.envs.get(e)
        covered = range(0, u + 1) if v == u + 1 else range(u, N)
        if ent is not None and all(psi.stamp[x] <= ent[1] for x in covered):
            return ent[0]
        if v == u + 1:
            t = env_left(self.environment(psi, (u - 1, u)), psi[u], self.H.W[u], self.H.c[u], self.H.c[u + 1], self.ctx)
        else:
            t = env_right(self.environment(psi, (u + 1, u)), psi[u], self.H.W[u], self.H.c[u], self.H.c[u + 1],
                          self.ctx)
        self.builds += 1
        self.envs[e] = (t, psi.clock)
        return t

    def position(self, psi: QNMPS, region):
        self.pos = sorted(region)
        lo, hi = self.pos[0], self.pos[-1]
        self.environment(psi, (lo - 1, lo))
        self.environment(psi, (hi + 1, hi))
        return self

    def left_right(self, psi):
        lo, hi = min(self.pos), max(self.pos)
        return self.environment(psi, (lo - 1, lo)), self.environment(psi, (hi + 1, hi))


# ------------------------------------------------------------ region step hooks
def extract(psi: QNMPS, region):
    """extracter.jl:3-8: the gauge centre walks into the region, then the local
    tensor is the site (1-site) or the product of the two sites (2-site)."""
    lo, hi = min(region), max(region)
    if len(region) not in (1, 2) or hi - lo != len(region) - 1:
        raise NotImplementedError("QN chains: 1-site regions and 2-site regions on neighbouring sites are served")
    orthogonalize(psi, region)
    psi.ortho_region = list(region)
    if len(region) == 1:
        return psi[lo].copy()
    return theta(psi[lo], psi[hi], psi.ctx)


def heff_map(P: QNProjTTN, psi: QNMPS, precision="fp64"):
    lo, hi = min(P.pos), max(P.pos)
    L, R = P.left_right(psi)
    if lo == hi:   # the block structure of the site, which the local tensor copies
        A = QNTensor(psi.ctx, psi[lo].indices, psi[lo].dirs, 0, keys=psi[lo].keys, zero=False)
        return heff1(L, P.H.W[lo], R, A, P.H.c[lo], P.H.c[lo + 1], psi.ctx, precision)
    th = QNTensor(psi.ctx, [psi.links[lo], psi.site, psi.site, psi.links[hi + 1]], [1, 1, 1, -1], 0,
                   zero=False)   # the block structure only
    return heff2(L, P.H.W[lo], P.H.W[hi], R, th, P.H.c[lo], P.H.c[hi + 1], psi.ctx, precision)


def insert(psi: QNMPS, region, th: QNTensor, trunc, which_decomp=None):
    """inserter.jl:3-33: a 1-site region stores the tensor (no truncation); a
    2-site region is split by the sector factorisation."""
    if len(region) == 1:
        psi[region[0]] = th
        return None
    u, v = region
    lo, hi = min(u, v), max(u, v)
    left, right, mid, info = factorize(th, u != lo, cutoff=trunc["cutoff"], maxdim=trunc["maxdim"],
                                       mindim=trunc["mindim"], which_decomp=which_decomp, ctx=psi.ctx)
    psi.links[lo + 1] = mid
    psi[lo] = left
    psi[hi] = right
    return info


# ------------------------------------------------------------ orthogonalize walk
_CHUNK = 24   # per-sector tensors per block program (a program addresses at most 30 buffers)


def _run_chunked(ctx, fixed, per_sector, emit):
    """Runs emit(program, fixed buffer ids, [(sector, buffer id)]) over chunks of
    the per-sector tensors so that every program stays within its buffer limit."""
    keys = list(per_sector)
    for c0 in range(0, max(len(keys), 1), _CHUNK):
        chunk = keys[c0:c0 + _CHUNK]
        if not chunk:
            break
        p = _Program()
        fid = [p.operand(t) for t in fixed]
        sid = [(q, p.operand(per_sector[q])) for q in chunk]
        emit(p, fid, sid)
        _run_once(p, ctx, None, None)


def qr_toward(psi: QNMPS, x, y):
    """One step of the orthogonalize walk (extracter.jl:6) on a QN chain: the gauge
    centre moves from site x to its neighbour y by the QR of a QN tensor, which is
    block-wise -- each charge of the bond between x and y is decomposed on its own
    (``nsmi_qr``, thin, non-negative diag(R)), so the bond keeps min(rows, cols)
    vectors per charge; x becomes the isometry and y absorbs R."""
    ctx = psi.ctx
    A, B = psi[x], psi[y]
    site = psi.site
    right = y == x + 1
    bond = x + 1 if right else x
    bix = psi.links[bond]
    if right:
        lx = A.indices[0]
        groups = _charge_rows(lx, site, 1, 1)           # rows (l, s) per charge
    else:
        rx = A.indices[2]
        groups = _charge_rows(site, rx, -1, 1)          # columns (s, r) per charge
    secs = [(b, bix.q[b]) for b in range(len(bix)) if bix.q[b] in groups]
    mats = {}
    for b, q in secs:
        n = groups[q][2]
        mats[q] = ctx.empty((n, bix.n[b]) if right else (bix.n[b], n))

    def gather(p, fid, sid):
        xA = fid[0]
        zero, cp = [], []
        for q, yq in sid:
            keys, off, n = groups[q]
            b = bix.find(q)
            zero.append((yq, 0, n * bix.n[b], []))
            for (i, j) in keys:
                key = (i, j, b) if right else (b, i, j)
                if not A.has(key):
                    continue
                if right:      # block l.n[i] x bond.n[b] at row off
                    cp.append((xA, A.block_off(key), bix.n[b], yq, off[(i, j)] * bix.n[b], bix.n[b], lx.n[i],
                               bix.n[b]))
                else:          # block bond.n[b] x r.n[j] at column off
                    cp.append((xA, A.block_off(key), rx.n[j], yq, off[(i, j)], n, bix.n[b], rx.n[j]))
        p.lincomb(zero)
        if cp:
            p.copy(cp)

    _run_chunked(ctx, [A.data], mats, gather)
    Qs, Rs, newsec = {}, {}, []
    for b, q in secs:
        Q, R = kernels.qr(mats[q], 1, not right)
        Qs[q], Rs[q] = Q, R
        newsec.append((q, Q.shape[1] if right else Q.shape[0]))
    m = QNIndex(newsec)
    if right:
        newA = QNTensor(ctx, [lx, site, m], [1, 1, -1], 0, zero=False)
        newB = QNTensor(ctx, [m] + B.indices[1:], [1, 1, -1], 0, zero=False)
    else:
        newA = QNTensor(ctx, [m, site, rx], [1, 1, -1], 0, zero=False)
        newB = QNTensor(ctx, B.indices[:2] + [m], [1, 1, -1], 0, zero=False)

    def scatter(p, fid, sid):
        yA = fid[0]
        cp = []
        for q, xq in sid:
            keys, off, n = groups[q]
            mm = m.find(q)
            k = m.n[mm]
            for (i, j) in keys:
                key = (i, j, mm) if right else (mm, i, j)
                if right:      # Q rows off .. + l.n[i], k columns
                    cp.append((xq, off[(i, j)] * k, k, yA, newA.block_off(key), k, lx.n[i], k))
                else:          # Q (k x n) columns off .. + r.n[j]
                    cp.append((xq, off[(i, j)], n, yA, newA.block_off(key), rx.n[j], k, rx.n[j]))
        p.copy(cp)

    _run_chunked(ctx, [newA.data], Qs, scatter)

    def absorb(p, fid, sid):
        xB, yB = fid
        items = []
        for q, xr in sid:
            mm, b = m.find(q), bix.find(q)
            k, nb = m.n[mm], bix.n[b]
            for key in newB.keys:
                if (key[0] if right else key[2]) != mm:
                    continue
                shp = newB.shapes[newB.pos[key]]
                okey = ((b,) + key[1:]) if right else (key[:2] + (b,))
                sg = []
                if right:      # R (k x nb) . B block (nb x w)
                    w = math.prod(shp[1:])
                    if B.has(okey):
                        sg.append((xr, 0, nb, xB, B.block_off(okey), w, nb, 1.0))
                    items.append((yB, newB.block_off(key), w, k, w, sg))
                else:          # B block (h x nb) . R (nb x k)
                    h = math.prod(shp[:-1])
                    if B.has(okey):
                        sg.append((xB, B.block_off(okey), nb, xr, 0, k, nb, 1.0))
                    items.append((yB, newB.block_off(key), k, h, k, sg))
        if items:
            p.gemm(False, False, items)

    _run_chunked(ctx, [B.data, newB.data], Rs, absorb)
    psi.links[bond] = m
    psi[x] = newA
    psi[y] = newB
    psi.ortho_region = [y]


def orthogonalize(psi: QNMPS, region):
    """Walks the gauge centre into the region (extracter.jl:6) by block-wise QR
    steps; a no-op when it is already there."""
    lo, hi = min(region), max(region)
    cur = psi.ortho_region
    if cur is not None and all(lo <= v <= hi for v in cur):
        return psi
    if cur is None or len(cur) != 1:
        raise NotImplementedError("QN chains: the gauge centre must be one site to walk it")
    c = cur[0]
    while not (lo <= c <= hi):
        nxt = c + 1 if c < lo else c - 1
        qr_toward(psi, c, nxt)
        c = nxt
    return psi


# ------------------------------------------------------------ subspace expansion
def _charge_rows(ix1: QNIndex, ix2: QNIndex, sign1, sign2):
    """Row blocks (i, j) of a matrix over two indices grouped by the charge
    sign1 q1[i] + sign2 q2[j], in (i, j) order: {charge: ([(i, j)], {(i, j): row
    offset}, rows)}; every block has ix1.n[i] * ix2.n[j] entries."""
    out = {}
    for i in range(len(ix1)):
        for j in range(len(ix2)):
            q = sign1 * ix1.q[i] + sign2 * ix2.q[j]
            keys, ro, n = out.setdefault(q, ([], {}, 0))
            keys.append((i, j))
            ro[(i, j)] = n
            out[q] = (keys, ro, n + ix1.n[i] * ix2.n[j])
    return out


def densitymatrix_expand(problem, local_state: QNTensor, region, *, trunc, north_pass=1, expansion_factor=1.5,
                         max_expand=None, **kws):
    """subspace/densitymatrix.jl:5-74 on a QN chain: the link between the
    operator's previous position p and the 2-site region grows by the dominant
    eigenvectors of rho = S S^T, S = (1 - A A^T)(E_p . W_p . A_p).  S is block
    diagonal in the charge Q of that link, so one block program forms S sector by
    sector (the environment update's GEMM and MPO-mix steps, then the projection
    as grouped GEMMs), ``nsmi_factorize_sectors`` diagonalises the sector Gram
    matrices with the eigen route and truncates across sectors (maxdim = the
    expansion size of subspace.jl:28-48), and a second program projects U off A
    again, forms the check A^T U, A <- [A | U], expander = [A | U]^T A, and applies
    the expander to the next site and the local tensor.  In every charge the new
    link lists the old basis first, then that charge's new vectors.  When
    |A^T U| > 1e-10 the reference prints a warning and changes nothing; so does
    this."""
    from .subspace import compute_expansion
    psi, P = problem.state, problem.operator
    lo, hi = min(region), max(region)
    prev = [v for v in P.sites() if v not in (lo, hi)]
    if len(prev) != 1 or prev[0] not in (lo - 1, hi + 1):
        return problem, local_state
    p = prev[0]
    left = p == lo - 1
    v = lo if left else hi
    ctx = psi.ctx
    A, C = psi[p], psi[v]
    site = psi.site
    a = A.indices[2] if left else A.indices[0]
    nexp = compute_expansion(a.dim, A.indices[0 if left else 2].dim * site.dim, expansion_factor=expansion_factor,
                             max_expand=max_expand, maxdim=trunc.get("maxdim"))
    if nexp <= 0:
        return problem, local_state
    if left:
        lx, _, a = A.indices                      # A[l, s, a]; rows (l, s)
        rows = _charge_rows(lx, site, 1, 1)
        E = P.environment(psi, (p - 1, p))
        mix = _mix_left(E, A, P.H.W[p], P.H.c[p])
        # Y block (ap, sp, wp, b): row (ap, sp), column (b, wp)
        yrow = lambda k: (k[0], k[1])             # noqa: E731
        ycol = lambda k: (k[3], k[2])             # noqa: E731
        ablk = lambda i, j, b: (i, j, b)          # noqa: E731  A's block of row (i, j) in a-sector b
    else:
        a, _, rx = A.indices                      # A[a, s, r]; rows (s, r)
        rows = _charge_rows(site, rx, -1, 1)
        E = P.environment(psi, (p + 1, p))
        mix = _mix_right(E, A, P.H.W[p], P.H.c[p + 1])
        # Y block (a, w, sp, bp): row (sp, bp), column (a, w)
        yrow = lambda k: (k[2], k[3])             # noqa: E731
        ycol = lambda k: (k[0], k[1])             # noqa: E731
        ablk = lambda i, j, b: (b, i, j)          # noqa: E731
    # sector layout: rows of every charge, columns (and so a matrix) where Y has blocks
    cols = {}
    for k in mix.Y:
        q = (lx.q[k[0]] + site.q[k[1]]) if left else (rx.q[k[3]] - site.q[k[2]])
        cols.setdefault(q, {})[ycol(k)] = None
    colw = (lambda c: a.n[c[0]])                  # noqa: E731  width of column block (b, w)
    sectors = sorted(q for q in cols if q in rows)
    co, C_ = {}, {}
    for q in sectors:
        off = 0
        for c in sorted(cols[q]):
            co[(q, c)] = off
            off += colw(c)
        C_[q] = off
    na = {q: (a.n[a.find(q)] if a.find(q) is not None else 0) for q in set(rows) | set(a.q)}
    aq = sorted(q for q in a.q if q in rows)       # charges of A's columns (A_Q matrices)
    M_off, tot = {}, 0
    for q in sectors:
        M_off[q] = tot
        tot += rows[q][2] * C_[q]
    A_off, totA = {}, 0
    for q in aq:
        A_off[q] = totA
        totA += rows[q][2] * na[q]
    Mall = ctx.empty((max(tot, 1),))
    Aall = ctx.empty((max(totA, 1),))
    # ---- program 1: S by sectors, A_Q, projection S_Q -= A_Q (A_Q^T S_Q)
    # left:  M_Q = S_Q  (rows x cols),  A_Q stored rows x na
    # right: M_Q = S_Q^T (cols x rows), A_Q stored transposed, na x rows
    p1 = _Program()
    xE, xA, yM, yA = p1.operand(E.data), p1.operand(A.data), p1.operand(Mall), p1.operand(Aall)
    sY = mix.emit(p1, xE, xA)
    proj = [q for q in sectors if na[q] > 0]
    x2off, zoff, t = {}, {}, 0
    for q in proj:
        x2off[q] = t
        t += na[q] * C_[q]
    j0 = len(p1.scratch)
    p1.scratch_buf(t)
    p1.scratch_buf(tot)
    sX2, sZ = p1.sid(j0), p1.sid(j0 + 1)
    zero = [(yM, M_off[q], rows[q][2] * C_[q], []) for q in sectors]
    zero += [(yA, A_off[q], rows[q][2] * na[q], []) for q in aq]
    if zero:
        p1.lincomb(zero)
    cp = []
    for k, yo in mix.Y.items():
        rk, ck = yrow(k), ycol(k)
        q = (lx.q[k[0]] + site.q[k[1]]) if left else (rx.q[k[3]] - site.q[k[2]])
        R, ro = rows[q][2], rows[q][1][rk]
        if left:   # Y block l.n[ap] x a.n[b] at (ro, co)
            cp.append((sY, yo, a.n[k[3]], yM, M_off[q] + ro * C_[q] + co[(q, ck)], C_[q], lx.n[k[0]], a.n[k[3]]))
        else:      # Y block a.n[a] x r.n[bp] at (co, ro) of the transposed layout
            cp.append((sY, yo, rx.n[k[3]], yM, M_off[q] + co[(q, ck)] * R + ro, R, a.n[k[0]], rx.n[k[3]]))
    for q in aq:
        b = a.find(q)
        R = rows[q][2]
        for (i, j) in rows[q][0]:
            key = ablk(i, j, b)
            if not A.has(key):
                continue
            ro = rows[q][1][(i, j)]
            if left:   # A block l.n[i] x na (site j) at rows ro
                cp.append((xA, A.block_off(key), na[q], yA, A_off[q] + ro * na[q], na[q], lx.n[i], na[q]))
            else:      # A block na x r.n[j] (site i) at columns ro of the na x R layout
                cp.append((xA, A.block_off(key), rx.n[j], yA, A_off[q] + ro, R, na[q], rx.n[j]))
    if cp:
        p1.copy(cp)
    for _ in range(north_pass if proj else 0):
        if left:
            p1.gemm(True, False, [(sX2, x2off[q], C_[q], na[q], C_[q],
                                   [(yA, A_off[q], na[q], yM, M_off[q], C_[q], rows[q][2], 1.0)]) for q in proj])
            p1.gemm(False, False, [(sZ, M_off[q], C_[q], rows[q][2], C_[q],
                                    [(yA, A_off[q], na[q], sX2, x2off[q], C_[q], na[q], 1.0)]) for q in proj])
        else:
            R = {q: rows[q][2] for q in proj}
            p1.gemm(False, True, [(sX2, x2off[q], na[q], C_[q], na[q],
                                   [(yM, M_off[q], R[q], yA, A_off[q], R[q], R[q], 1.0)]) for q in proj])
            p1.gemm(False, False, [(sZ, M_off[q], R[q], C_[q], R[q],
                                    [(sX2, x2off[q], na[q], yA, A_off[q], R[q], na[q], 1.0)]) for q in proj])
        p1.lincomb([(yM, M_off[q], rows[q][2] * C_[q], [(yM, M_off[q], 1.0), (sZ, M_off[q], -1.0)])
                    for q in proj])
    _run_once(p1, ctx, None, None)
    # ---- the truncated eigen-decomposition of rho by sectors
    Rl = [rows[q][2] for q in sectors]
    Cl = [C_[q] for q in sectors]
    Uall, _, ks, _ = _factorize_sectors(ctx, Mall, [M_off[q] for q in sectors], Rl if left else Cl,
                                        Cl if left else Rl, not left, trunc.get("cutoff", 0.0), nexp,
                                        trunc.get("mindim", 1), "eigen")
    kq = {q: k for q, k in zip(sectors, ks) if k > 0}
    if not kq:
        return problem, local_state
    U_off, t = {}, 0
    for q, k in zip(sectors, ks):
        U_off[q] = t
        t += k * rows[q][2]
    # ---- program 2: U projection, the check, [A | U], the expander and its products
    # left: U_Q rows x k;  right: V_Q = U_Q^T, k x rows
    new_q = sorted(set(a.q) | set(kq))
    ax = QNIndex([(q, na.get(q, 0) + kq.get(q, 0)) for q in new_q])
    if left:
        newA = QNTensor(ctx, [lx, site, ax], [1, 1, -1], 0, zero=False)
        newC = QNTensor(ctx, [ax] + C.indices[1:], [1, 1, -1], 0, zero=False)
        newT = QNTensor(ctx, [ax] + local_state.indices[1:], local_state.dirs, 0, zero=False)
    else:
        newA = QNTensor(ctx, [ax, site, rx], [1, 1, -1], 0, zero=False)
        newC = QNTensor(ctx, C.indices[:2] + [ax], [1, 1, -1], 0, zero=False)
        newT = QNTensor(ctx, local_state.indices[:-1] + [ax], local_state.dirs, 0, zero=False)
    both = [q for q in kq if na[q] > 0]            # charges where U meets A
    c_off, t = {}, 0
    for q in both:
        c_off[q] = t
        t += na[q] * kq[q]
    chk = ctx.empty((max(t, 1),))
    ex_q = [q for q in new_q if na.get(q, 0) > 0]   # expander blocks (n_ax x na)
    ex_off, te = {}, 0
    for q in ex_q:
        ex_off[q] = te
        te += (na[q] + kq.get(q, 0)) * na[q]
    p2 = _Program()
    xAa, xU, xA2, xC, xT = p2.operand(Aall), p2.operand(Uall), p2.operand(A.data), p2.operand(C.data), \
        p2.operand(local_state.data)
    yNA, yNC, yNT, yK = p2.operand(newA.data), p2.operand(newC.data), p2.operand(newT.data), p2.operand(chk)
    j0 = len(p2.scratch)
    p2.scratch_buf(t)
    p2.scratch_buf(sum(kq[q] * rows[q][2] for q in both))
    p2.scratch_buf(te)
    sX3, sZ3, sEx = p2.sid(j0), p2.sid(j0 + 1), p2.sid(j0 + 2)
    z_off, tz = {}, 0
    for q in both:
        z_off[q] = tz
        tz += kq[q] * rows[q][2]

    def at_u(out_buf, out_off):   # A_Q^T U_Q (left) or V_Q A_Q (right): na x k / k x na
        if left:
            return [(out_buf, out_off[q], kq[q], na[q], kq[q],
                     [(xAa, A_off[q], na[q], xU, U_off[q], kq[q], rows[q][2], 1.0)]) for q in both]
        return [(out_buf, out_off[q], na[q], kq[q], na[q],
                 [(xU, U_off[q], rows[q][2], xAa, A_off[q], rows[q][2], rows[q][2], 1.0)]) for q in both]

    for _ in range(north_pass if both else 0):
        if left:
            p2.gemm(True, False, at_u(sX3, c_off))
            p2.gemm(False, False, [(sZ3, z_off[q], kq[q], rows[q][2], kq[q],
                                    [(xAa, A_off[q], na[q], sX3, c_off[q], kq[q], na[q], 1.0)]) for q in both])
        else:
            p2.gemm(False, True, at_u(sX3, c_off))
            p2.gemm(False, False, [(sZ3, z_off[q], rows[q][2], kq[q], rows[q][2],
                                    [(sX3, c_off[q], na[q], xAa, A_off[q], rows[q][2], na[q], 1.0)]) for q in both])
        p2.lincomb([(xU, U_off[q], kq[q] * rows[q][2], [(xU, U_off[q], 1.0), (sZ3, z_off[q], -1.0)])
                    for q in both])
    if both:
        p2.gemm(left, not left, at_u(yK, c_off))
    # [A | U]: zero, then A's blocks and U's rows / columns
    p2.lincomb([(yNA, 0, newA.size, [])])
    cp = []
    for key in newA.keys:
        m = key[2] if left else key[0]
        q = ax.q[m]
        i, j = (key[0], key[1]) if left else (key[1], key[2])
        n_old, k = na.get(q, 0), kq.get(q, 0)
        nax = ax.n[m]
        b = a.find(q)
        if n_old and A.has(ablk(i, j, b)):
            src = A.block_off(ablk(i, j, b))
            if left:
                cp.append((xA2, src, n_old, yNA, newA.block_off(key), nax, lx.n[i], n_old))
            else:
                cp.append((xA2, src, rx.n[j], yNA, newA.block_off(key), rx.n[j], n_old, rx.n[j]))
        if k:
            ro = rows[q][1][(i, j)]
            if left:
                cp.append((xU, U_off[q] + ro * k, k, yNA, newA.block_off(key) + n_old, nax, lx.n[i], k))
            else:
                R = rows[q][2]
                cp.append((xU, U_off[q] + ro, R, yNA, newA.block_off(key) + n_old * rx.n[j], rx.n[j], k, rx.n[j]))
    if cp:
        p2.copy(cp)
    # expander E_Q = [A_Q | U_Q]^T A_Q: na x na on top, then k x na
    items = []
    for q in ex_q:
        R, n_old, k = rows[q][2], na[q], kq.get(q, 0)
        if left:
            items.append((sEx, ex_off[q], n_old, n_old, n_old, [(xAa, A_off[q], n_old, xAa, A_off[q], n_old, R, 1.0)]))
            if k:
                items.append((sEx, ex_off[q] + n_old * n_old, n_old, k, n_old,
                              [(xU, U_off[q], k, xAa, A_off[q], n_old, R, 1.0)]))
    if items:
        p2.gemm(True, False, items)
    items = []
    for q in ex_q:
        R, n_old, k = rows[q][2], na[q], kq.get(q, 0)
        if not left:
            items.append((sEx, ex_off[q], n_old, n_old, n_old, [(xAa, A_off[q], R, xAa, A_off[q], R, R, 1.0)]))
            if k:
                items.append((sEx, ex_off[q] + n_old * n_old, n_old, k, n_old,
                              [(xU, U_off[q], R, xAa, A_off[q], R, R, 1.0)]))
    if items:
        p2.gemm(False, True, items)
    # next site and local tensor: E . C on the left link, or C . E^T on the right link
    for T_old, T_new, y_new, x_old in ((C, newC, yNC, xC), (local_state, newT, yNT, xT)):
        items = []
        for key in T_new.keys:
            m = key[0] if left else key[-1]
            q = ax.q[m]
            nax, n_old = ax.n[m], na.get(q, 0)
            b = a.find(q)
            okey = ((b,) + key[1:]) if left else (key[:-1] + (b,))
            sg = []
            shp = T_new.shapes[T_new.pos[key]]
            if left:
                w = math.prod(shp[1:])
                if n_old and T_old.has(okey):
                    sg.append((sEx, ex_off[q], n_old, x_old, T_old.block_off(okey), w, n_old, 1.0))
                items.append((y_new, T_new.block_off(key), w, nax, w, sg))
            else:
                h = math.prod(shp[:-1])
                if n_old and T_old.has(okey):
                    sg.append((x_old, T_old.block_off(okey), n_old, sEx, ex_off[q], n_old, n_old, 1.0))
                items.append((y_new, T_new.block_off(key), nax, h, nax, sg))
        if items:
            if left:
                p2.gemm(False, False, items)
            else:
                p2.gemm(False, True, items)
    _run_once(p2, ctx, None, None)
    if both:
        err = chk.norm()
        if err > 1e-10:
            print(f"Warning: |U*A| = {err:.3E} in subspace expansion")
            return problem, local_state
    ia = p + 1 if left else p
    psi.links[ia] = ax
    psi[p] = newA
    psi[v] = newC
    problem.state = psi
    return problem, newT


def expect_mpo(psi: QNMPS, H: QNMPO) -> float:
    """<psi|H|psi> / <psi|psi> by a left-to-right environment pass (and one with the
    identity operator for the norm)."""
    ctx = psi.ctx
    E = _boundary_env(ctx, psi.links[0], H.links[0])
    for i in range(len(psi)):
        E = env_left(E, psi[i], H.W[i], H.c[i], H.c[i + 1], ctx)
    d = psi.site.dim
    one = np.eye(d).reshape(1, 1, d, d)
    z = np.zeros(1, dtype=np.int64)
    N1 = _boundary_env(ctx, psi.links[0], QNIndex([(0, 1)]))
    for i in range(len(psi)):
        N1 = env_left(N1, psi[i], one, z, z, ctx)
    return float(E.data.numpy().sum()) / float(N1.data.numpy().sum())


__all__ = ["QNIndex", "QNTensor", "QNMPO", "QNMPS", "QNProjTTN", "spin_site_index", "qn_product_mps", "theta",
           "env_left", "env_right", "heff1", "heff2", "factorize", "expect_mpo", "BlockProgram", "densitymatrix_expand",
           "qr_toward", "orthogonalize"]
