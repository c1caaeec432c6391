"""Sweep driver mirroring NetworkSolvers.jl's iterator API on top of the device kernels.

    eigsolve / dmrg      eigsolve.jl:45-76
    EigsolveProblem      eigsolve.jl:4-12
    sweep_iterator       iterators.jl:5-40      (SweepIterator)
    region_iterator      iterators.jl:46-79     (RegionIterator)
    region_iterator_action  iterators.jl:81-100 (extracter -> updater -> inserter)
    region_tuples        adapters.jl:12-32
    sweep_solve          sweep_solve.jl:12-40
    extracter            extracter.jl:3-17
    inserter             inserter.jl:3-33
    optimal_map          operator_map.jl:3-10
    eigsolve_solver      local_solvers/eigsolve.jl:3-29

Keyword packs keep the reference's names: ``extracter_kwargs``,
``updater_kwargs``, ``inserter_kwargs`` and ``trunc=dict(cutoff, maxdim,
mindim)``, with per-sweep vector schedules.  Added keywords:
``precision`` (updater: arithmetic of the H_eff contractions, "fp64" | "fp32"),
``which_decomp`` (inserter: force "svd" / "eigen").
"""
from __future__ import annotations

import functools
import itertools
from dataclasses import dataclass, field

from . import kernels, qn, subspace, tree
from .network import ProjTTN, TTN, TTNO, orthogonalize
from .region_plans import euler_sweep, truncation_parameters


# ------------------------------------------------------------------- problems
@dataclass
class EigsolveProblem:
    """eigsolve.jl:4-12."""
    state: TTN
    operator: ProjTTN
    eigenvalue: float = float("inf")
    region_log: list = field(default_factory=list)


def state(p):
    return p.state


def operator(p):
    return p.operator


def eigenvalue(p):
    return p.eigenvalue


# ------------------------------------------------------------------ iterators
class RegionIterator:
    """iterators.jl:46-71; iterating performs one region step per element."""

    def __init__(self, problem, region_plan):
        self.problem = problem
        self.region_plan = region_plan
        self.which_region = 0

    def current_region_plan(self):
        return self.region_plan[self.which_region]

    def current_region(self):
        return self.current_region_plan()[0]

    def region_kwargs(self):
        return self.current_region_plan()[1]

    def previous_region(self):
        return None if self.which_region == 0 else self.region_plan[self.which_region - 1][0]

    def next_region(self):
        return None if self.which_region == len(self.region_plan) - 1 else \
            self.region_plan[self.which_region + 1][0]

    def is_last_region(self):
        return self.next_region() is None

    def __iter__(self):
        for which in range(len(self.region_plan)):
            self.which_region = which
            _, kw = self.region_plan[which]
            self.problem = region_iterator_action(self.problem, self, **kw)
            yield self


def current_region(R: RegionIterator):
    return R.current_region()


def next_region(R: RegionIterator):
    return R.next_region()


@functools.singledispatch
def region_plan(problem, **kws):
    """iterators.jl:102-104: Euler-tour plan for eigensolver problems."""
    return euler_sweep(state(problem).graph, **kws)


def region_iterator(problem, **sweep_kwargs):
    """iterators.jl:77-79."""
    return RegionIterator(problem, region_plan(problem, **sweep_kwargs))


def region_iterator_action(problem, region_iter, *, extracter_kwargs=None, updater_kwargs=None,
                           inserter_kwargs=None, sweep, **kws):
    """iterators.jl:81-100: extracter -> updater -> inserter with kwarg routing."""
    extracter_kwargs = extracter_kwargs or {}
    updater_kwargs = updater_kwargs or {}
    inserter_kwargs = inserter_kwargs or {}
    # later keyword packs win, as Julia's splatted keywords merge
    problem, local_state = extracter(problem, region_iter, **{**extracter_kwargs, "sweep": sweep, **kws})
    problem, local_state = updater(problem, local_state, region_iter, **{**updater_kwargs, **kws})
    problem = inserter(problem, local_state, region_iter, **{"sweep": sweep, **inserter_kwargs, **kws})
    return problem


class SweepIterator:
    """iterators.jl:5-31: rebuilds the region iterator with the sweep's kwargs."""

    def __init__(self, sweep_kws, region_iter):
        self.sweep_kws = list(sweep_kws)
        self.region_iter = region_iter
        self.which_sweep = 1

    def __len__(self):
        return len(self.sweep_kws)

    def problem(self):
        return self.region_iter.problem

    def __iter__(self):
        for i, kw in enumerate(self.sweep_kws):
            if i > 0:
                self.region_iter = region_iterator(self.region_iter.problem, sweep=self.which_sweep, **kw)
            self.which_sweep += 1
            yield self.region_iter


def sweep_iterator(problem, sweep_kws=None, nsweeps=None, **kws):
    """iterators.jl:33-40: either an explicit list of per-sweep kwargs, or
    ``nsweeps`` repetitions of ``kws``."""
    if sweep_kws is None:
        sweep_kws = list(itertools.repeat(kws, nsweeps))
    sweep_kws = list(sweep_kws)
    return SweepIterator(sweep_kws, region_iterator(problem, sweep=1, **sweep_kws[0]))


def region_tuples(R: RegionIterator):
    """adapters.jl:18-32: yields (region, region_kwargs) after each region step."""
    for r in R:
        yield r.current_region(), r.region_kwargs()


def default_region_callback(problem, **kws):
    return None


def default_sweep_callback(problem, **kws):
    return None


def default_sweep_printer(problem, *, outputlevel, sweep, nsweeps, **kws):
    if outputlevel >= 1:
        print(f"Done with sweep {sweep}/{nsweeps}")


def sweep_solve(sweep_iter: SweepIterator, *, outputlevel=0, region_callback=default_region_callback,
                sweep_callback=default_sweep_callback, sweep_printer=default_sweep_printer, **kwargs):
    """sweep_solve.jl:12-40."""
    nsweeps = len(sweep_iter)
    for sweep, region_iter in enumerate(sweep_iter, start=1):
        for region, region_kwargs in region_tuples(region_iter):
            region_callback(region_iter.problem, nsweeps=nsweeps, outputlevel=outputlevel, region=region,
                            region_kwargs=region_kwargs, sweep=sweep, **kwargs)
        sweep_callback(region_iter, nsweeps=nsweeps, outputlevel=outputlevel, sweep=sweep, **kwargs)
        sweep_printer(region_iter, nsweeps=nsweeps, outputlevel=outputlevel, sweep=sweep, **kwargs)
    return sweep_iter.problem()


# -------------------------------------------------------------- region phases
def extracter(problem, region_iter, *, sweep, trunc=None, subspace_algorithm=None, **kws):
    """extracter.jl:3-17: orthogonalize to the region, theta = product of the region
    tensors, subspace expansion (``subspace_algorithm``: None or "densitymatrix",
    subspace/densitymatrix.jl), then ``position`` builds the environments."""
    t = truncation_parameters(sweep, **(trunc or {}))
    region = current_region(region_iter)
    if subspace_algorithm not in (None, "densitymatrix"):
        raise NotImplementedError(
            "Subspace expansion (subspace_expand!) not defined for requested combination of "
            f"subspace_algorithm={subspace_algorithm!r} and problem types")
    if isinstance(state(problem), qn.QNMPS):
        local_state = qn.extract(state(problem), list(region))
        problem, local_state = subspace.subspace_expand(problem, local_state, region,
                                                        subspace_algorithm=subspace_algorithm, trunc=t, **kws)
        problem.operator = operator(problem).position(problem.state, list(region))
        return problem, local_state
    if isinstance(state(problem), tree.TreeTTN):
        psi = state(problem)
        local_state = tree.extract(psi, list(region))
        problem, local_state = subspace.subspace_expand(problem, local_state, region,
                                                        subspace_algorithm=subspace_algorithm, trunc=t, **kws)
        problem.operator = operator(problem).position(problem.state, list(region))
        return problem, local_state
    psi = orthogonalize(state(problem), region)
    problem.state = psi
    lo, hi = min(region), max(region)
    if len(region) == 2:
        if hi != lo + 1:
            raise ValueError(f"2-site region {region} is not a chain bond")
        local_state = kernels.theta(psi[lo], psi[hi])
    elif len(region) == 1:
        local_state = psi[lo].copy()
    else:
        raise ValueError(f"Region of length {len(region)} not currently supported")
    problem, local_state = subspace.subspace_expand(problem, local_state, region,
                                                    subspace_algorithm=subspace_algorithm, trunc=t, **kws)
    problem.operator = operator(problem).position(problem.state, region)
    return problem, local_state


def optimal_map(P: ProjTTN, psi: TTN, precision="fp64") -> kernels.HeffMap:
    """operator_map.jl:3-10: the region's effective Hamiltonian as a device map
    (environments of the incident edges and the region's MPO tensors; on a tree a
    contraction program over up to four environments)."""
    if isinstance(P, qn.QNProjTTN):
        return qn.heff_map(P, psi, precision)
    if isinstance(P, tree.TreeProjTTN):
        region = P.sites()
        return P.netmap(psi, tree.region_labels(psi, region), tree.region_shape(psi, region), precision)
    L, R = P.left_right(psi)
    sites = P.sites()
    H = P.operator()
    if len(sites) == 2:
        return kernels.HeffMap(L, H[sites[0]], H[sites[1]], R, precision)
    return kernels.HeffMap(L, H[sites[0]], None, R, precision)


def eigsolve_solver(operator_map: kernels.HeffMap, init, howmany=1, *, which_eigval="SR", ishermitian=True,
                    tol=1e-14, krylovdim=3, maxiter=1, verbosity=0, eager=False, **kws):
    """local_solvers/eigsolve.jl:3-29: lowest Ritz pair by the device Lanczos
    (KrylovKit eigsolve semantics); returns (eigenvalue, local_state)."""
    if howmany != 1 or which_eigval != "SR" or not ishermitian or eager:
        raise NotImplementedError("eigsolve_solver: only howmany=1, :SR, ishermitian, eager=false are served")
    x = init.copy()
    res = kernels.lanczos_eig(operator_map, x, krylovdim=krylovdim, maxiter=maxiter, tol=tol)
    return res.eigval, x


@functools.singledispatch
def updater(problem, local_state, region_iter, **kws):
    raise NotImplementedError(f"no updater for {type(problem).__name__}")


@updater.register
def _(E: EigsolveProblem, local_state, region_iter, *, outputlevel=0, solver=eigsolve_solver,
      precision="fp64", **kws):
    """eigsolve.jl:14-28."""
    H = optimal_map(operator(E), state(E), precision)
    try:
        eigval, local_state = solver(H, local_state, **kws)
    finally:
        H.close()
    E.eigenvalue = eigval
    if outputlevel >= 2:
        print(f"  Region {current_region(region_iter)}: energy = {eigval:.12f}")
    return E, local_state


def inserter(problem, local_tensor, region_iter, *, normalize=False, set_orthogonal_region=True, sweep,
             trunc=None, which_decomp=None, **kws):
    """inserter.jl:3-33: truncated factorize with U on first(region) and C on
    last(region); 1-site regions store the tensor; longer regions raise."""
    t = truncation_parameters(sweep, **(trunc or {}))
    region = current_region(region_iter)
    psi = state(problem)
    info = None
    if isinstance(psi, qn.QNMPS):
        if len(region) > 2:
            raise ValueError(f"Region of length {len(region)} not currently supported")
        info = qn.insert(psi, list(region), local_tensor, t, which_decomp)
    elif isinstance(psi, tree.TreeTTN):
        if len(region) > 2:
            raise ValueError(f"Region of length {len(region)} not currently supported")
        info = tree.insert(psi, list(region), local_tensor, t, which_decomp)
    else:
        if len(region) == 1:
            C = local_tensor
        elif len(region) == 2:
            u, v = region
            lo = min(u, v)
            U, C, info = kernels.factorize(local_tensor, 2, u != lo, cutoff=t["cutoff"], maxdim=t["maxdim"],
                                           mindim=t["mindim"], which_decomp=which_decomp)
            psi[u] = U
        else:
            raise ValueError(f"Region of length {len(region)} not currently supported")
        psi[region[-1]] = C
    v = region[-1]
    if set_orthogonal_region:
        psi.ortho_region = [v]
    if normalize:
        psi[v].scale_(1.0 / psi[v].norm())
    if info is not None and hasattr(problem, "region_log"):
        problem.region_log.append((tuple(region), getattr(problem, "eigenvalue", None), info.kept,
                                   info.truncerr, info.route))
    problem.state = psi
    return problem


# ------------------------------------------------------------------- top level
def eigsolve_sweep_printer(region_iter, *, outputlevel, sweep, nsweeps, **kws):
    """eigsolve.jl:30-43."""
    if outputlevel >= 1:
        E = region_iter.problem
        head = f"After sweep {sweep:02d}/{nsweeps} " if nsweeps >= 10 else f"After sweep {sweep}/{nsweeps} "
        print(head + f"eigenvalue={eigenvalue(E):.12f} maxlinkdim={state(E).maxlinkdim()}", flush=True)


def eigsolve_problem(init_prob: EigsolveProblem, *, nsweeps, nsites=1, outputlevel=0, extracter_kwargs=None,
                     updater_kwargs=None, inserter_kwargs=None, sweep_printer=eigsolve_sweep_printer, **kws):
    """eigsolve.jl:45-67 (problem form)."""
    sweep_iter = sweep_iterator(init_prob, nsweeps=nsweeps, nsites=nsites, outputlevel=outputlevel,
                                extracter_kwargs=extracter_kwargs or {}, updater_kwargs=updater_kwargs or {},
                                inserter_kwargs=inserter_kwargs or {})
    prob = sweep_solve(sweep_iter, outputlevel=outputlevel, sweep_printer=sweep_printer, **kws)
    return eigenvalue(prob), state(prob)


def eigsolve(H: TTNO, init_state: TTN, *, env_precision="fp64", **kws):
    """eigsolve.jl:69-74: wraps (H, psi0) into an EigsolveProblem.  The state is
    copied; ``env_precision`` selects the environment-update arithmetic."""
    init_prob = EigsolveProblem(state=init_state.copy(), operator=projected_operator(H, env_precision))
    return eigsolve_problem(init_prob, **kws)


def projected_operator(H, env_precision="fp64"):
    """ProjTTN of a chain MPO or a tree TTN-MPO (eigsolve.jl:71, applyexp.jl:86)."""
    if isinstance(H, tree.TreeTTNO):
        return tree.TreeProjTTN(H, env_precision)
    if isinstance(H, qn.QNMPO):
        return qn.QNProjTTN(H, H.ctx)
    return ProjTTN(H, env_precision)


def dmrg(*args, **kws):
    """eigsolve.jl:76."""
    return eigsolve(*args, **kws)
