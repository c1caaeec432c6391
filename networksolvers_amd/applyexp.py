"""Time evolution by TDVP: exp(t H) applied to a chain state, mirroring applyexp.jl.

    ApplyExpProblem       applyexp.jl:4-12
    region_plan           applyexp.jl:14-16 -> tdvp_regions (tdvp_region_plans.jl:34-44)
    updater               applyexp.jl:18-48 (local solver, 1-site bond back-evolution)
    applyexp / tdvp       applyexp.jl:62-103
    exponentiate_solver   local_solvers/exponentiate.jl:3-29 (device Krylov exponential)
    runge_kutta_solver    local_solvers/runge_kutta.jl:31-54

The local solvers receive the region's device effective Hamiltonian (a
``kernels.HeffMap``) where the reference passes ``x -> optimal_map(P, x)``.
Complex time steps need a complex state: ``applyexp`` promotes a real initial
state to complex128 when any exponent is complex.
"""
from __future__ import annotations

import functools
from dataclasses import dataclass, field

import numpy as np

from . import kernels, tree
from .network import ProjTTN, TTN, TTNO, _absorb_left, _absorb_right
from .region_plans import tdvp_regions
from .sweep import (current_region, next_region, operator, optimal_map, projected_operator, region_plan, state,
                    sweep_iterator, sweep_solve, updater)


@dataclass
class ApplyExpProblem:
    """applyexp.jl:4-12."""
    state: TTN
    operator: ProjTTN
    current_time: complex = 0.0
    region_log: list = field(default_factory=list)


def current_time(T: ApplyExpProblem):
    return T.current_time


@region_plan.register
def _(T: ApplyExpProblem, *, nsites, time_step, **sweep_kwargs):
    """applyexp.jl:14-16."""
    return tdvp_regions(state(T).graph, time_step, nsites=nsites, **sweep_kwargs)


# -------------------------------------------------------------- local solvers
def exponentiate_solver(operator_map: kernels.HeffMap, time, init, *, krylovdim=30, maxiter=100, verbosity=0,
                        tol=1e-12, ishermitian=True, issymmetric=True, eager=True, **kws):
    """local_solvers/exponentiate.jl:3-29: exp(time H) init by the device Krylov
    exponential (Lanczos basis, error per unit time <= tol); returns (x, info)."""
    if not ishermitian:
        raise NotImplementedError("exponentiate_solver: only Hermitian operators are served")
    x = init.copy()
    info = kernels.krylov_expm(operator_map, x, time, krylovdim=krylovdim, maxiter=maxiter, tol=tol)
    return x, info


def _runge_kutta_2(H, t, psi0):
    """local_solvers/runge_kutta.jl:31-35."""
    Hpsi = H(psi0)
    H2psi = H(Hpsi)
    out = psi0.copy()
    out.axpy_(t, Hpsi)
    out.axpy_(t * t / 2, H2psi)
    return out


def _runge_kutta_4(H, t, psi0):
    """local_solvers/runge_kutta.jl:37-43, with every stage a device H_eff apply
    and device axpy:  k1 = H psi0, k2 = k1 + t/2 H k1, k3 = k1 + t/2 H k2,
    k4 = k1 + t H k3, psi0 + t/6 (k1 + 2 k2 + 2 k3 + k4)."""
    k1 = H(psi0)
    k2 = H(k1).scale_(t / 2).axpy_(1.0, k1)
    k3 = H(k2).scale_(t / 2).axpy_(1.0, k1)
    k4 = H(k3).scale_(t).axpy_(1.0, k1)
    out = psi0.copy()
    out.axpy_(t / 6, k1).axpy_(t / 3, k2).axpy_(t / 3, k3).axpy_(t / 6, k4)
    return out


def runge_kutta_solver(H, time, psi, *, order=4, **kws):
    """local_solvers/runge_kutta.jl:45-54."""
    if order == 4:
        return _runge_kutta_4(H, time, psi), {}
    if order == 2:
        return _runge_kutta_2(H, time, psi), {}
    raise ValueError("For runge_kutta_solver, must specify `order` keyword")


# -------------------------------------------------------------------- updater
_bond_mpo_cache: dict = {}


def _identity_mpo(ctx, D):
    """W[w, w', 1, 1] = delta(w, w'): with it the 1-site H_eff of a bond tensor
    [l, 1, r] is the bond effective Hamiltonian sum_w L[:, w, :] x R[:, w, :]^T."""
    key = (id(ctx), D)
    t = _bond_mpo_cache.get(key)
    if t is None or t.ctx is not ctx or ctx.closed:
        t = ctx.upload(np.eye(D).reshape(D, D, 1, 1))
        _bond_mpo_cache[key] = t
    return t


def _back_evolve_bond(T: ApplyExpProblem, v, forward, x, time_step, solver, precision, kws):
    """applyexp.jl:30-42 on a chain: x = Q Rb over the bond from v to the next
    region's vertex (QR with non-negative diag), Rb evolved by -time_step with
    the bond effective Hamiltonian (the environment of Q's side rebuilt with Q,
    the other side's environment as cached), returns Q Rb(t)."""
    P = operator(T)
    psi = state(T)
    L, R = P.left_right(psi)
    W = P.operator()[v]
    if forward:
        Q, Rb = kernels.qr(x, 2, False)                            # x[(l s), r] = Q[(l s), k] Rb[k, r]
        Lb, Rbd = kernels.env_update(L, Q, W, 0, P.env_precision), R
        bond = Rb.reshape_((Rb.shape[0], 1, Rb.shape[1]))
    else:
        Q, Rb = kernels.qr(x, 1, True)                             # x[l, (s r)] = Rb[l, k] Q[k, (s r)]
        Lb, Rbd = L, kernels.env_update(R, Q, W, 1, P.env_precision)
        bond = Rb.reshape_((Rb.shape[0], 1, Rb.shape[1]))
    Hb = kernels.HeffMap(Lb, _identity_mpo(P.ctx, Lb.shape[1]), None, Rbd, precision)
    try:
        bt, _ = solver(Hb, -time_step, bond, **kws)
    finally:
        Hb.close()
    bt.reshape_((bt.shape[0], bt.shape[2]))
    if forward:
        return _absorb_right(Q, bt)                                # Q[l, s, k] bt[k, r]
    return _absorb_left(bt, Q)                                     # bt[l, k] Q[k, s, r]


@updater.register
def _(T: ApplyExpProblem, local_state, region_iter, *, nsites, time_step, solver=runge_kutta_solver,
      outputlevel=0, precision="fp64", **kws):
    """applyexp.jl:18-48: evolve the region tensor by time_step with the local
    solver; for 1-site sweeps, when the next region differs, evolve the bond to
    it backwards; then advance current_time by time_step (as the reference, every
    region adds its own step, so 1-site sweeps advance N steps per sweep)."""
    H = optimal_map(operator(T), state(T), precision)
    try:
        local_state, _ = solver(H, time_step, local_state, **kws)
    finally:
        H.close()
    if nsites == 1:
        curr, nxt = current_region(region_iter), next_region(region_iter)
        if nxt is not None and list(nxt) != list(curr):
            (v,) = curr
            if isinstance(state(T), tree.TreeTTN):
                local_state = tree.back_evolve_bond(T, v, list(nxt), local_state, time_step, solver, kws, precision)
            else:
                local_state = _back_evolve_bond(T, v, nxt[0] > v, local_state, time_step, solver, precision, kws)
    T.current_time = T.current_time + time_step
    if outputlevel >= 2:
        print(f"  Region {current_region(region_iter)}: time step {time_step}")
    return T, local_state


# ------------------------------------------------------------------ top level
def applyexp_sweep_printer(region_iter, *, outputlevel, sweep, nsweeps, process_time=lambda t: t, **kws):
    """applyexp.jl:50-60."""
    if outputlevel >= 1:
        T = region_iter.problem
        print(f"  Current time = {process_time(current_time(T))}, maxlinkdim={state(T).maxlinkdim()}", flush=True)


def applyexp_problem(init_prob: ApplyExpProblem, exponents, *, extracter_kwargs=None, updater_kwargs=None,
                     inserter_kwargs=None, outputlevel=0, nsites=1, tdvp_order=4,
                     sweep_printer=applyexp_sweep_printer, **kws):
    """applyexp.jl:62-82: one sweep per exponent increment,
    time_steps = diff([0, exponents...])[2:end]."""
    pts = [0.0] + list(exponents)
    time_steps = [pts[i + 1] - pts[i] for i in range(len(pts) - 1)][1:]
    if not time_steps:
        return state(init_prob)
    sweep_kws = dict(outputlevel=outputlevel, extracter_kwargs=extracter_kwargs or {},
                     inserter_kwargs=inserter_kwargs or {}, nsites=nsites, tdvp_order=tdvp_order,
                     updater_kwargs=updater_kwargs or {})
    kws_array = [dict(sweep_kws, time_step=t) for t in time_steps]
    sweep_iter = sweep_iterator(init_prob, kws_array)
    prob = sweep_solve(sweep_iter, outputlevel=outputlevel, sweep_printer=sweep_printer, **kws)
    return prob


def _as_complex(psi):
    if isinstance(psi, tree.TreeTTN):
        return tree.TreeTTN.from_arrays(psi.ctx, psi.tg, psi.to_arrays(), dtype=np.complex128)
    return TTN.from_arrays(psi.ctx, psi.to_arrays(), psi.graph, dtype=np.complex128)


def applyexp(H: TTNO, init_state: TTN, exponents, *, env_precision="fp64", **kws) -> TTN:
    """applyexp.jl:84-89: exp(z H) psi0 sampled at the given exponents z.  The
    state is copied (promoted to complex128 if any exponent is complex)."""
    cplx = any(isinstance(z, complex) and z.imag != 0 for z in exponents) or np.iscomplexobj(np.asarray(exponents))
    psi = _as_complex(init_state) if (cplx and init_state.dtype != np.complex128) else init_state.copy()
    psi.ortho_region = init_state.ortho_region
    init_prob = ApplyExpProblem(state=psi, operator=projected_operator(H, env_precision))
    return state(applyexp_problem(init_prob, exponents, **kws))


def process_real_times(z):
    """applyexp.jl:91."""
    return round(-complex(z).imag, 10)


def tdvp(H: TTNO, init_state: TTN, time_points, *, process_time=process_real_times, sweep_printer=None, **kws):
    """applyexp.jl:93-103: real-time evolution, exponents = -i t."""
    if sweep_printer is None:
        sweep_printer = functools.partial(applyexp_sweep_printer, process_time=process_time)
    exponents = [-1j * t for t in time_points]
    return applyexp(H, init_state, exponents, sweep_printer=sweep_printer, **kws)


__all__ = ["ApplyExpProblem", "current_time", "exponentiate_solver", "runge_kutta_solver", "applyexp",
           "applyexp_problem", "tdvp", "process_real_times", "applyexp_sweep_printer"]
