"""MPOs of the benchmark Hamiltonians (test infrastructure only).

The reference builds these with ITensorMPS ``OpSum`` -> ``itn.mpo``/``itn.ttn``
(test/dmrg/test_tree_dmrg.jl:21-28, examples/dmrg.jl:11-18); here they are
written as the standard lower-triangular finite-state-machine MPO.  The MPO
gauge changes rounding only.  Pinned by: dense MPO vs explicit Kronecker sums
and ED energies (tests/test_oracle_chain.py).
"""
from __future__ import annotations

import numpy as np


def spin_ops(d):
    """(Sz, S+, S-, I) for spin S=(d-1)/2 in the basis m = S, S-1, ..., -S
    ("Up" first, as the ITensors S=1/2 and S=1 site types)."""
    S = (d - 1) / 2
    m = S - np.arange(d)
    Sz = np.diag(m)
    Sp = np.zeros((d, d))
    for k in range(1, d):
        # <m+1| S+ |m> = sqrt(S(S+1) - m(m+1)), row k-1 has m = m[k]+1
        Sp[k - 1, k] = np.sqrt(S * (S + 1) - m[k] * (m[k] + 1))
    return Sz, Sp, Sp.T.copy(), np.eye(d)


def xxz_mpo(N, *, d=2, Jxy=1.0, Jz=1.0, hz=0.0):
    """H = sum_i [Jxy/2 (S+_i S-_{i+1} + S-_i S+_{i+1}) + Jz Sz_i Sz_{i+1}] + hz sum_i Sz_i,
    open boundaries, D = 5:
        W = [[I, 0, 0, 0, 0], [S+, 0, 0, 0, 0], [S-, 0, 0, 0, 0], [Sz, 0, 0, 0, 0],
             [hz Sz, Jxy/2 S-, Jxy/2 S+, Jz Sz, I]]   (rows: left bond, cols: right bond)
    W[0] keeps row 4 and W[N-1] keeps column 0.  Layout W[wl, wr, s', s]."""
    Sz, Sp, Sm, I = spin_ops(d)
    D = 5
    W = np.zeros((D, D, d, d))
    W[0, 0] = I
    W[1, 0] = Sp
    W[2, 0] = Sm
    W[3, 0] = Sz
    W[4, 0] = hz * Sz
    W[4, 1] = 0.5 * Jxy * Sm
    W[4, 2] = 0.5 * Jxy * Sp
    W[4, 3] = Jz * Sz
    W[4, 4] = I
    mpo = []
    for i in range(N):
        w = W
        if i == 0:
            w = w[4:5]
        if i == N - 1:
            w = w[:, 0:1]
        mpo.append(np.ascontiguousarray(w))
    return mpo


def heisenberg_mpo(N, d=2, J=1.0):
    return xxz_mpo(N, d=d, Jxy=J, Jz=J)


def tfim_mpo(N, *, h=1.0, J=1.0):
    """Chain TFIM with Pauli matrices: H = -J sum Z_i Z_{i+1} - h sum X_i (D = 3)."""
    X = np.array([[0.0, 1.0], [1.0, 0.0]])
    Z = np.array([[1.0, 0.0], [0.0, -1.0]])
    I = np.eye(2)
    W = np.zeros((3, 3, 2, 2))
    W[0, 0] = I
    W[1, 0] = Z
    W[2, 0] = -h * X
    W[2, 1] = -J * Z
    W[2, 2] = I
    mpo = []
    for i in range(N):
        w = W
        if i == 0:
            w = w[2:3]
        if i == N - 1:
            w = w[:, 0:1]
        mpo.append(np.ascontiguousarray(w))
    return mpo
