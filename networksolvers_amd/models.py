"""Input formats of the benchmark workloads: spin-chain MPOs and seeded MPS arrays.

The reference builds its operators with ITensorMPS ``OpSum`` -> ``itn.mpo`` and
its states with ``itn.random_mps`` / ``itn.ttn(state, sites)``
(examples/dmrg.jl:11-24, test/dmrg/test_tree_dmrg.jl:21-35); here they are
numpy arrays in the device layouts (W[wl, wr, s', s], A[l, s, r]) that
``TTNO``/``TTN`` upload.
"""
from __future__ import annotations

import numpy as np


def spin_operators(d: int):
    """(Sz, S+, S-, I) for spin (d-1)/2 in the basis m = S, ..., -S."""
    S = (d - 1) / 2
    m = S - np.arange(d)
    Sz = np.diag(m)
    Sp = np.zeros((d, d))
    for k in range(1, d):
        Sp[k - 1, k] = np.sqrt(S * (S + 1) - m[k] * (m[k] + 1))
    return Sz, Sp, Sp.T.copy(), np.eye(d)


def xxz_mpo(N: int, *, d: int = 2, Jxy: float = 1.0, Jz: float = 1.0, hz: float = 0.0):
    """Open-chain XXZ MPO (D = 5), H = sum Jxy/2 (S+S- + S-S+) + Jz SzSz + hz sum Sz.
    Finite-state machine rows: left bond, columns: right bond; W[0] keeps the
    last row, W[N-1] the first column."""
    Sz, Sp, Sm, I = spin_operators(d)
    W = np.zeros((5, 5, d, d))
    W[0, 0], W[1, 0], W[2, 0], W[3, 0] = I, Sp, Sm, Sz
    W[4, 0], W[4, 1], W[4, 2], W[4, 3], W[4, 4] = hz * Sz, 0.5 * Jxy * Sm, 0.5 * Jxy * Sp, Jz * Sz, I
    out = []
    for i in range(N):
        w = W[4:5] if i == 0 else W
        w = w[:, 0:1] if i == N - 1 else w
        out.append(np.ascontiguousarray(w))
    return out


def heisenberg_mpo(N: int, d: int = 2, J: float = 1.0):
    return xxz_mpo(N, d=d, Jxy=J, Jz=J)


def random_mps(N: int, d: int, chi: int, seed: int):
    """Seeded MPS: bond dims min(chi, d^k, d^(N-k)), i.i.d. N(0,1) entries from
    numpy PCG64(seed) drawn site by site, normalised to one."""
    rng = np.random.Generator(np.random.PCG64(seed))
    dims = [1] + [int(min(chi, d ** k, d ** (N - k))) for k in range(1, N)] + [1]
    mps = [rng.standard_normal((dims[i], d, dims[i + 1])) for i in range(N)]
    E = np.ones((1, 1))
    for a in mps:
        E = np.tensordot(a, np.tensordot(E, a, axes=([1], [0])), axes=([0, 1], [0, 1]))
    mps[0] = mps[0] / np.sqrt(E[0, 0])
    return mps


def product_mps(states, d: int):
    """Product state with site i in basis state states[i]."""
    out = []
    for s in states:
        a = np.zeros((1, d, 1))
        a[0, s, 0] = 1.0
        out.append(a)
    return out
