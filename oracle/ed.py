"""Exact diagonalisation of spin Hamiltonians on arbitrary graphs (test infrastructure only).

Plays the role of the reference's brute-force ``ed_ground_state``
(test/utilities/simple_ed_methods.jl:4-16) for systems too large for dense
exponentials: the Hamiltonian is assembled sparse in a fixed total-Sz sector and
its lowest eigenvalue found with scipy's Lanczos (ARPACK ``eigsh``).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as sla


def _sector_states(N, d, m2_total):
    """Basis states (base-d digit arrays, digit k <-> m = S - k) with 2*sum(m) = m2_total."""
    S2 = d - 1  # 2S
    codes = np.arange(d ** N, dtype=np.int64)
    digits = np.empty((len(codes), N), dtype=np.int8)
    c = codes.copy()
    for i in range(N - 1, -1, -1):
        digits[:, i] = c % d
        c //= d
    m2 = (S2 - 2 * digits.astype(np.int64)).sum(axis=1)
    keep = m2 == m2_total
    return codes[keep], digits[keep]


def xxz_hamiltonian(N, edges, *, d=2, Jxy=1.0, Jz=1.0, m2_total=0):
    """Sparse H = sum_<ij> [Jxy/2 (S+_i S-_j + S-_i S+_j) + Jz Sz_i Sz_j] in the sector
    2*Sz_total = m2_total.  Returns (H, codes)."""
    codes, digits = _sector_states(N, d, m2_total)
    dim = len(codes)
    S = (d - 1) / 2
    m = S - digits.astype(np.float64)
    pw = d ** np.arange(N - 1, -1, -1, dtype=np.int64)
    diag = np.zeros(dim)
    rows, cols, vals = [], [], []
    for (i, j) in edges:
        diag += Jz * m[:, i] * m[:, j]
        # S+_i S-_j: digit_i -> digit_i - 1 (m_i + 1), digit_j -> digit_j + 1 (m_j - 1)
        for (a, b) in ((i, j), (j, i)):
            ok = (digits[:, a] > 0) & (digits[:, b] < d - 1)
            src = np.nonzero(ok)[0]
            ma, mb = m[src, a], m[src, b]
            amp = np.sqrt(S * (S + 1) - ma * (ma + 1)) * np.sqrt(S * (S + 1) - mb * (mb - 1))
            new = codes[src] - pw[a] + pw[b]
            dst = np.searchsorted(codes, new)
            rows.append(dst)
            cols.append(src)
            vals.append(0.5 * Jxy * amp)
    rows.append(np.arange(dim))
    cols.append(np.arange(dim))
    vals.append(diag)
    H = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(dim, dim))
    return H, codes


def ground_energy(H, k=1, tol=1e-13):
    if H.shape[0] <= 64:
        return float(np.linalg.eigvalsh(H.toarray())[0])
    w = sla.eigsh(H, k=k, which="SA", tol=tol, return_eigenvectors=False)
    return float(np.min(w))


def chain_edges(N):
    return [(i, i + 1) for i in range(N - 1)]


def xxz_chain_ground_energy(N, *, d=2, Jxy=1.0, Jz=1.0):
    """Lowest energy over all total-Sz sectors' relevant one: the Sz = 0 (or 1/2)
    sector holds the ground state of the antiferromagnetic chains used here."""
    m2 = 0 if (N * (d - 1)) % 2 == 0 else 1
    H, _ = xxz_hamiltonian(N, chain_edges(N), d=d, Jxy=Jxy, Jz=Jz, m2_total=m2)
    return ground_energy(H)
