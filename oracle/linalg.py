"""Dense FP64 linear-algebra restatements used by the region step (test infrastructure only).

Restated third-party behaviour (packages absent from this container; see each function):
  * NDTensors ``truncate!``           -- spectrum truncation rule (inserter.jl:23 via ITensors)
  * ITensors ``factorize``            -- svd/eigen route choice + truncated split (inserter.jl:23)
  * ITensors ``qr``                   -- gauge walk used by orthogonalize (extracter.jl:6, applyexp.jl:37)
  * KrylovKit ``eigsolve`` (Lanczos)  -- local_solvers/eigsolve.jl:16
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

SVD_EIGEN_THRESHOLD = 1e-12  # ITensors factorize: cutoff <= 1e-12 -> "svd", else "eigen"


# ---------------------------------------------------------------------------
# gauge conventions shared with the device path
# ---------------------------------------------------------------------------
def fix_column_phases(U: np.ndarray, C: np.ndarray | None = None):
    """Make the largest-magnitude entry of every column of U real and positive.

    The first index attaining the maximum magnitude is used.  The inverse phase
    is applied to the matching row of C so that U @ C is unchanged.  This is a
    gauge choice (the reference leaves the LAPACK sign); the device path applies
    the same rule so tensors can be compared element-wise.
    """
    if U.shape[1] == 0:
        return U, C
    idx = np.argmax(np.abs(U), axis=0)
    piv = U[idx, np.arange(U.shape[1])]
    mag = np.abs(piv)
    ph = np.where(mag > 0, piv / np.where(mag > 0, mag, 1.0), 1.0)
    U = U / ph[None, :]
    if C is not None:
        C = C * ph[:, None]
    return U, C


# ---------------------------------------------------------------------------
# NDTensors truncate!
# ---------------------------------------------------------------------------
def truncate(P, *, mindim=1, maxdim=None, cutoff=0.0,
             use_absolute_cutoff=False, use_relative_cutoff=True):
    """Restates NDTensors ``truncate!(P; mindim, maxdim, cutoff, ...)`` (called by
    ITensors ``svd``/``eigen`` inside ``factorize``, reached from inserter.jl:23).

    P is a spectrum sorted by decreasing magnitude (s^2 for svd, eigenvalues for
    eigen).  Returns (n_kept, truncerr, docut).  Steps, in the package's order:
    length-1 spectra are never truncated; a negative leading sign is flipped;
    the negative tail is zeroed; entries beyond maxdim are dropped (accumulating
    the discarded weight); then, with the relative cutoff, entries are dropped
    while ``truncerr + P[n] <= cutoff * sum(P)`` and ``n > mindim``; truncerr is
    divided by sum(P); n is clamped to >= 1; docut is the midpoint of the last
    kept and first dropped values (nudged by 1e-3*P[n] when nearly degenerate).

    Pinned by: tests/test_oracle_linalg.py hand-computed spectra (maxdim clamp,
    cutoff accumulation, mindim floor, negative tail, length-1 case).
    """
    P = np.array(P, dtype=np.float64, copy=True)
    origm = len(P)
    if maxdim is None:
        maxdim = origm
    if origm == 0:
        return 0, 0.0, 0.0
    if origm == 1:
        return 1, 0.0, abs(P[0]) / 2
    if P[0] < 0:
        P *= -1.0
    for n in range(origm - 1, -1, -1):
        if P[n] >= 0:
            break
        P[n] = 0.0
    n = origm  # 1-based count of kept entries
    truncerr = 0.0
    while n > maxdim:
        truncerr += P[n - 1]
        n -= 1
    if use_absolute_cutoff:
        while n > mindim and P[n - 1] <= cutoff:
            truncerr += P[n - 1]
            n -= 1
    else:
        scale = 1.0
        if use_relative_cutoff:
            scale = float(np.sum(P))
            if scale == 0.0:
                scale = 1.0
        while n > mindim and (truncerr + P[n - 1] <= cutoff * scale):
            truncerr += P[n - 1]
            n -= 1
        truncerr /= scale
    if n < 1:
        n = 1
    docut = 0.0
    if n < origm:
        docut = (P[n - 1] + P[n]) / 2
        if abs(P[n - 1] - P[n]) < 1e-3 * P[n - 1]:
            docut += 1e-3 * P[n - 1]
    return n, float(truncerr), float(docut)


# ---------------------------------------------------------------------------
# ITensors factorize (ortho="left")
# ---------------------------------------------------------------------------
@dataclass
class FactorizeInfo:
    route: str            # "svd" or "eigen"
    spectrum: np.ndarray  # full sorted spectrum (s^2 or eigenvalues)
    kept: int
    truncerr: float


def choose_route(cutoff, which_decomp=None, threshold=SVD_EIGEN_THRESHOLD):
    """ITensors ``factorize`` route choice: an explicit ``which_decomp`` wins;
    otherwise "svd" when ``cutoff is None or cutoff <= 1e-12`` and "eigen" when
    ``cutoff > 1e-12`` (the reference always passes a cutoff:
    truncation_parameters.jl:3,10).  The inclusive comparison at exactly 1e-12
    is a recalled detail of ITensors (parity unpinned); it is a parameter here."""
    if which_decomp is not None:
        return which_decomp
    if cutoff is None or cutoff <= threshold:
        return "svd"
    return "eigen"


def factorize(M: np.ndarray, *, cutoff=0.0, maxdim=None, mindim=1,
              which_decomp=None, threshold=SVD_EIGEN_THRESHOLD):
    """Truncated left-orthogonal split M = U @ C (ITensors ``factorize`` with
    ortho="left", as called at inserter.jl:23; U goes on first(region), C on
    last(region), inserter.jl:24,28-29).

    * maxdim is clamped to min(rows, cols) (ITensors factorize).
    * "svd": P = s^2 from the SVD of M, U = left singular vectors, C = S V^H.
    * "eigen": rho = M M^H, Hermitian eigendecomposition, eigenvalues sorted by
      decreasing |lambda| (NDTensors eigen), U = kept eigenvectors, C = U^H M.
    * truncation via ``truncate`` above; columns of U get ``fix_column_phases``.

    Pinned by: numpy svd/eigh reconstruction identities, U^H U = I,
    ||M - U C||^2 = truncerr * ||M||^2 (tests/test_oracle_linalg.py).
    """
    M = np.asarray(M)
    m, n = M.shape
    mx = min(m, n) if maxdim is None else min(int(maxdim), min(m, n))
    route = choose_route(cutoff, which_decomp, threshold)
    if route == "svd":
        u, s, vh = np.linalg.svd(M, full_matrices=False)
        P = s.astype(np.float64) ** 2
        k, truncerr, _ = truncate(P, mindim=mindim, maxdim=mx, cutoff=cutoff)
        U = u[:, :k]
        C = s[:k, None] * vh[:k, :]
    elif route == "eigen":
        rho = M @ M.conj().T
        rho = (rho + rho.conj().T) / 2  # LAPACK reads one triangle; symmetrise explicitly
        w, v = np.linalg.eigh(rho)
        order = np.argsort(-np.abs(w), kind="stable")
        P = w[order].astype(np.float64)
        v = v[:, order]
        k, truncerr, _ = truncate(P, mindim=mindim, maxdim=mx, cutoff=cutoff)
        U = v[:, :k]
        C = U.conj().T @ M
    else:
        raise ValueError(f"unknown decomposition {route!r}")
    U, C = fix_column_phases(U, C)
    return U, C, FactorizeInfo(route, P, k, truncerr)


# ---------------------------------------------------------------------------
# QR with a positive-diagonal convention
# ---------------------------------------------------------------------------
def qr_pos(M: np.ndarray):
    """Thin QR, M = Q R with r = min(m, n) and real non-negative diag(R).

    Restates the ITensors ``qr`` used by the orthogonalize walk (extracter.jl:6)
    and 1-site TDVP (applyexp.jl:37); the diagonal sign convention is a gauge
    choice made here (the device path uses the same), so Q and R are unique for
    full-rank input.  Pinned by: Q^H Q = I, Q R = M, R upper triangular."""
    Q, R = np.linalg.qr(M, mode="reduced")
    d = np.diagonal(R).copy()
    mag = np.abs(d)
    ph = np.where(mag > 0, d / np.where(mag > 0, mag, 1.0), 1.0)
    Q = Q * ph[None, :]
    R = R / ph[:, None]
    return Q, R


# ---------------------------------------------------------------------------
# small symmetric tridiagonal eigenproblem
# ---------------------------------------------------------------------------
def tridiag_eigh(alpha, beta):
    """Eigenpairs of the symmetric tridiagonal matrix with diagonal alpha and
    off-diagonal beta, ascending eigenvalues; each eigenvector's first
    non-negligible component is made positive (a gauge fixing the Ritz vector's
    overlap with the starting vector to be >= 0; the device path uses the same)."""
    K = len(alpha)
    T = np.diag(np.asarray(alpha, dtype=np.float64))
    for j in range(K - 1):
        T[j, j + 1] = T[j + 1, j] = beta[j]
    w, y = np.linalg.eigh(T)
    for j in range(K):
        col = y[:, j]
        i0 = int(np.argmax(np.abs(col) > 1e-14 * np.max(np.abs(col))))
        if col[i0] < 0:
            y[:, j] = -col
    return w, y


# ---------------------------------------------------------------------------
# KrylovKit eigsolve (Lanczos, ishermitian=true, which=:SR)
# ---------------------------------------------------------------------------
@dataclass
class LanczosInfo:
    numops: int
    krylov_dim: int
    alphas: list = field(default_factory=list)
    betas: list = field(default_factory=list)
    normres: float = 0.0


def _inner(a, b):
    return np.vdot(a, b)


def lanczos_eigsolve(apply, x0, *, krylovdim=3, maxiter=1, tol=1e-14,
                     eta=1 / math.sqrt(2), howmany=1):
    """Lowest Ritz pair of a Hermitian linear map from a Lanczos factorisation.

    Restates ``KrylovKit.eigsolve(f, x0, 1, :SR; ishermitian=true, tol,
    krylovdim, maxiter, eager=false)`` as called by eigsolve_solver
    (local_solvers/eigsolve.jl:3-29) with KrylovKit's default orthogonaliser
    ModifiedGramSchmidtIR(eta=1/sqrt(2)):

      initialize:  beta0=|x0|, Ax0=f(x0), alpha=<x0,Ax0>/beta0^2, v=x0/beta0,
                   r=Ax0/beta0 - alpha v, reorthogonalise r against v while
                   eps < |r| < eta*|r_old|.
      expand:      v_k = r/|r|; w=f(v_k); w -= beta_{k-1} v_{k-1}; alpha=<v_k,w>;
                   w -= alpha v_k; while eps < |w| < eta*sqrt(|w|^2+alpha^2+beta^2):
                   full MGS of w against the basis, alpha += last coefficient.
      stop:        at K == krylovdim or |r| <= tol (invariant subspace).
      result:      lowest eigenpair (D[0], y) of the K x K tridiagonal matrix
                   (``tridiag_eigh`` gauge), Ritz vector = sum_j y_j v_j.

    Only maxiter == 1 is restated (the reference default and every configured
    case).  Pinned by: exactness when dim <= krylovdim, comparison with
    scipy's eigh_tridiagonal on random symmetric matrices, the invariant-subspace
    stop on rank-deficient operators (tests/test_oracle_linalg.py).
    """
    if maxiter != 1:
        raise NotImplementedError("only maxiter=1 (KrylovKit without restart) is restated")
    x0 = np.asarray(x0)
    beta0 = float(np.linalg.norm(x0))
    if beta0 == 0.0:
        raise ValueError("initial vector should not have norm zero")
    Ax0 = apply(x0)
    numops = 1
    alpha = (_inner(x0, Ax0) / (beta0 * beta0)).real
    v = x0 / beta0
    r = Ax0 / beta0
    beta_old = float(np.linalg.norm(r))
    r = r - alpha * v
    beta = float(np.linalg.norm(r))
    epsv = np.finfo(np.float64).eps
    while epsv < beta < eta * beta_old:
        beta_old = beta
        da = _inner(v, r)
        alpha += da.real
        r = r - da * v   # the full (complex) projection is removed; alpha keeps its real part
        beta = float(np.linalg.norm(r))
    V = [v]
    alphas = [alpha]
    betas = []
    while True:
        K = len(V)
        if K == krylovdim or beta <= tol:
            break
        # expand
        v_new = r / beta
        betas.append(beta)
        vprev = V[-1]
        V.append(v_new)
        w = apply(v_new)
        numops += 1
        w = w - beta * vprev
        ca = _inner(v_new, w)
        w = w - ca * v_new
        a = ca.real
        ab2 = a * a + beta * beta
        b = float(np.linalg.norm(w))
        nold = math.sqrt(b * b + ab2)
        while epsv < b < eta * nold:
            nold = b
            for q in V:  # modified Gram-Schmidt against the whole basis
                c = _inner(q, w)
                w = w - c * q
            a += c.real  # coefficient on V[end]
            b = float(np.linalg.norm(w))
        alphas.append(a)
        beta = b
        r = w
    K = len(V)
    D, Y = tridiag_eigh(alphas, betas)
    y = Y[:, 0]
    x = sum(y[j] * V[j] for j in range(K))
    info = LanczosInfo(numops=numops, krylov_dim=K, alphas=list(alphas),
                       betas=list(betas), normres=abs(beta * y[-1]))
    return float(D[0]), x, info
