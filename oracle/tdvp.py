"""TDVP / applyexp on a chain (test infrastructure only).

Restates applyexp.jl:18-103 (ApplyExpProblem updater, applyexp, tdvp) with the
region plans of tdvp_region_plans.jl, the local solvers of
local_solvers/exponentiate.jl and local_solvers/runge_kutta.jl, and the extracter
/ inserter of the DMRG loop.
"""
from __future__ import annotations

import math

import numpy as np

from . import linalg
from .chain import ChainEnvs, env_left, env_right, heff1_apply, heff2_apply, orthogonalize_chain, theta2
from .plans import get_or_last, tdvp_regions


def _phi1(z):
    return 1.0 + z * (0.5 + z / 6.0) if abs(z) < 1e-5 else (np.exp(z) - 1.0) / z


def krylov_expm(apply, tau, x, *, krylovdim=30, maxiter=100, tol=1e-12, eta=1 / math.sqrt(2)):
    """exp(tau H) x for Hermitian H (exponentiate_solver, local_solvers/exponentiate.jl:3-29).

    Lanczos basis from the current vector (recurrence and ModifiedGramSchmidtIR
    re-orthogonalisation as in ``linalg.lanczos_eigsolve``).  After each new basis
    vector the projected exponential y = exp(z T_K) e1, z = sign(tau) dt, comes
    from the eigendecomposition of T_K, and the step is accepted when Saad's error
    estimate beta0 beta_K dt |[phi1(z T_K) e1]_K| <= tol dt (tol per unit time);
    at K = krylovdim the step shrinks by clamp(0.9 (tol dt / err)^(1/K), 0.1, 0.9)
    until the estimate passes, the accepted step size carries over to the next
    sub-step, and sub-steps continue until the full time is covered (at most
    maxiter).  KrylovKit's own
    phi-function step-size control is not restated (parity unpinned: it changes
    the result only within tol).  Pinned by scipy.linalg.expm (tests)."""
    total = abs(tau)
    w = np.asarray(x).copy()
    info = {"numops": 0, "substeps": 0, "kmax": 0, "err": 0.0}
    if total == 0:
        return w, info
    sgn = tau / total
    done = 0.0
    dt_next = total
    epsv = np.finfo(np.float64).eps
    while total - done > 1e-14 * total:
        if info["substeps"] >= maxiter:
            raise RuntimeError("exponentiate did not converge within maxiter sub-steps")
        info["substeps"] += 1
        beta0 = float(np.linalg.norm(w))
        if beta0 == 0:
            break
        Aw = apply(w)
        info["numops"] += 1
        alpha = (np.vdot(w, Aw) / (beta0 * beta0)).real
        v = w / beta0
        r = Aw / beta0
        beta_old = float(np.linalg.norm(r))
        r = r - alpha * v
        beta = float(np.linalg.norm(r))
        while epsv < beta < eta * beta_old:
            beta_old = beta
            da = np.vdot(v, r)
            alpha += da.real
            r = r - da * v
            beta = float(np.linalg.norm(r))
        V = [v]
        alphas, betas = [alpha], []
        dt = min(total - done, dt_next)
        dt_next = total
        while True:
            K = len(V)
            T = np.diag(alphas) + np.diag(betas, 1) + np.diag(betas, -1)
            lam, S = np.linalg.eigh(T)

            def evaluate(step):
                z = sgn * step
                y = S @ (np.exp(z * lam) * S[0])
                phik = S[K - 1] @ (np.array([_phi1(z * l) for l in lam]) * S[0])
                return y, beta0 * beta * step * abs(phik)

            y, err = evaluate(dt)
            if beta <= 1e-14 * beta0 or err <= tol * dt:
                break
            if K == krylovdim:
                tries = 0
                while err > tol * dt and tries < 60:
                    dt *= min(0.9, max(0.1, 0.9 * (tol * dt / err) ** (1.0 / K)))
                    y, err = evaluate(dt)
                    tries += 1
                dt_next = dt
                break
            # expand
            betas.append(beta)
            vprev = V[-1]
            vk = r / beta
            V.append(vk)
            wv = apply(vk)
            info["numops"] += 1
            wv = wv - beta * vprev
            ca = np.vdot(vk, wv)
            wv = wv - ca * vk
            a = ca.real
            b = float(np.linalg.norm(wv))
            nold = math.sqrt(b * b + a * a + beta * beta)
            while epsv < b < eta * nold:
                nold = b
                for q in V:
                    c = np.vdot(q, wv)
                    wv = wv - c * q
                a += c.real
                b = float(np.linalg.norm(wv))
            alphas.append(a)
            beta = b
            r = wv
        info["kmax"] = max(info["kmax"], len(V))
        info["err"] = err
        w = beta0 * sum(y[j] * V[j] for j in range(len(V)))
        done += dt
    return w, info


def runge_kutta_4(apply, t, x):
    """local_solvers/runge_kutta.jl:8-14."""
    k1 = apply(x)
    k2 = k1 + (t / 2) * apply(k1)
    k3 = k1 + (t / 2) * apply(k2)
    k4 = k1 + t * apply(k3)
    return x + (t / 6) * (k1 + 2 * k2 + 2 * k3 + k4)


def tdvp_chain(W, mps0, time_points, **kws):
    """tdvp (applyexp.jl:93-103): applyexp_chain with exponents -i t."""
    return applyexp_chain(W, mps0, [-1j * t for t in time_points], **kws)


def applyexp_chain(W, mps0, exponents, *, nsites=2, tdvp_order=4, solver="exponentiate", maxdim=None,
                   cutoff=0.0, mindim=1, normalize=False, krylovdim=30, tol=1e-12, root_vertex=None):
    """applyexp (applyexp.jl:62-89) on a chain: one sweep per time step (diff of
    [0, exponents...] dropping its first entry), region plan
    tdvp_regions (post-order plan from root_vertex, order-1/2/4 sub-steps, backward
    1-site steps between 2-site regions), updater = local solver with the region's
    time step (plus, for nsites=1, the QR and backward bond evolution of
    applyexp.jl:30-42), inserter = factorize with U on first(region), C on last.
    Returns (mps, current_time)."""
    N = len(W)
    root = N - 1 if root_vertex is None else root_vertex
    cplx = any(complex(z).imag != 0 for z in exponents) or any(np.iscomplexobj(a) for a in mps0)
    mps = [np.array(a, dtype=np.complex128 if cplx else np.float64) for a in mps0]
    envs = ChainEnvs(W)
    ortho = None
    pts = [0.0] + list(exponents)
    steps = [pts[i + 1] - pts[i] for i in range(len(pts) - 1)][1:]
    if not cplx:
        steps = [float(np.real(z)) for z in steps]
    vs, es = list(range(N)), [(i, i + 1) for i in range(N - 1)]
    current_time = 0.0

    def solve(apply, tau, x):
        if solver == "exponentiate":
            return krylov_expm(apply, tau, x, krylovdim=krylovdim, tol=tol)[0]
        return runge_kutta_4(apply, tau, x)

    for sweep, dt in enumerate(steps, start=1):
        cut = get_or_last(cutoff, sweep)
        mxd = get_or_last(maxdim, sweep) if maxdim is not None else None
        mnd = get_or_last(mindim, sweep)
        plan = tdvp_regions(vs, es, root, dt, tdvp_order, nsites)
        for idx, (region, tau) in enumerate(plan):
            lo, hi = min(region), max(region)
            if ortho is None or not (lo <= ortho <= hi):
                mps = orthogonalize_chain(mps, lo, hi)
                for i in range(N):
                    envs.site_changed(i)
            L = envs.get(mps, lo - 1, lo)
            R = envs.get(mps, hi + 1, hi)
            if len(region) == 2:
                th = theta2(mps[lo], mps[hi])
                shape = th.shape
                x = solve(lambda t: heff2_apply(L, W[lo], W[hi], R, t.reshape(shape)).reshape(-1), tau,
                          th.reshape(-1)).reshape(shape)
                cl, d1, d2, cr = shape
                M = x.reshape(cl * d1, d2 * cr)
                u, v = region
                if u == lo:
                    U, C, _ = linalg.factorize(M, cutoff=cut, maxdim=mxd, mindim=mnd)
                    mps[lo] = U.reshape(cl, d1, -1)
                    mps[hi] = C.reshape(-1, d2, cr)
                else:
                    U, C, _ = linalg.factorize(M.T, cutoff=cut, maxdim=mxd, mindim=mnd)
                    mps[hi] = U.T.reshape(-1, d2, cr)
                    mps[lo] = C.T.reshape(cl, d1, -1)
                v_last = v
            else:
                (v,) = region
                a = mps[v]
                shape = a.shape
                x = solve(lambda t: heff1_apply(L, W[v], R, t.reshape(shape)).reshape(-1), tau,
                          a.reshape(-1)).reshape(shape)
                if nsites == 1 and idx + 1 < len(plan) and plan[idx + 1][0] != region:
                    x = _back_evolve_bond(solve, L, R, W[v], x, tau, forward=plan[idx + 1][0][0] > v)
                mps[v] = x
                v_last = v
            if normalize:
                mps[v_last] = mps[v_last] / np.linalg.norm(mps[v_last])
            for i in set(region):
                envs.site_changed(i)
            ortho = v_last
            current_time += tau
    return mps, current_time


def _back_evolve_bond(solve, L, R, Wv, x, tau, forward):
    """applyexp.jl:30-42: split the updated site tensor x = Q Rb over the bond to the
    next region (QR with non-negative diag), evolve the bond tensor by -tau with
    the bond effective Hamiltonian (environments of both sides of the bond, Q
    absorbed into its side), and return Q Rb(t)."""
    cl, d, cr = x.shape
    if forward:
        Q, Rb = linalg.qr_pos(x.reshape(cl * d, cr))
        Q = Q.reshape(cl, d, -1)
        Lb, Rbond, bond = env_left(L, Q, Wv), R, Rb
    else:
        Q, Rb = linalg.qr_pos(x.reshape(cl, d * cr).T)
        Q = Q.T.reshape(-1, d, cr)
        Lb, Rbond, bond = L, env_right(R, Q, Wv), Rb.T
    D = Lb.shape[1]
    Wid = np.eye(D).reshape(D, D, 1, 1)
    shape = (bond.shape[0], 1, bond.shape[1])
    bt = solve(lambda t: heff1_apply(Lb, Wid, Rbond, t.reshape(shape)).reshape(-1), -tau,
               bond.reshape(-1)).reshape(bond.shape)
    if forward:
        return np.tensordot(Q, bt, axes=([2], [0]))
    return np.tensordot(bt, Q, axes=([1], [0]))
