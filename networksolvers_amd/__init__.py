"""networksolvers_amd -- MI355X-native region-step engine for NetworkSolvers.jl's sweeps.

The region step (two-site effective-Hamiltonian apply inside a Lanczos
eigensolver, environment updates, truncated factorisation) runs in
``libnsmi.so`` (HIP/CDNA4 kernels behind a C ABI, include/nsmi.h).  The Python
modules mirror the reference's iterator API, region plans and keyword packs.
Importing the package loads the library and raises if it is missing.
"""
from . import _lib

_lib.load()

from .device import Context, DeviceTensor  # noqa: E402
from .kernels import HeffMap, env_update, factorize, krylov_expm, lanczos_eig, qr, theta  # noqa: E402
from .models import heisenberg_mpo, product_mps, random_mps, xxz_mpo  # noqa: E402
from .network import ProjTTN, TTN, TTNO, expect_local, expect_mpo, orthogonalize  # noqa: E402
from .region_plans import (Graph, euler_sweep, euler_tour_edges, euler_tour_vertices, path_graph,  # noqa: E402
                           post_order_dfs_plan, tdvp_regions, truncation_parameters)
from .sweep import (EigsolveProblem, dmrg, eigsolve, eigsolve_solver, extracter, inserter,  # noqa: E402
                    optimal_map, region_iterator, sweep_iterator, sweep_solve, updater)
from .applyexp import (ApplyExpProblem, applyexp, current_time, exponentiate_solver,  # noqa: E402
                       process_real_times, runge_kutta_solver, tdvp)
from .kernels import NetMap, permute, tensordot  # noqa: E402
from .qn import QNIndex, QNMPO, QNMPS, QNProjTTN, QNTensor, qn_product_mps, spin_site_index  # noqa: E402
from .tree import (TreeGraph, TreeProjTTN, TreeTTN, TreeTTNO, heisenberg_tree_mpo, product_ttn,  # noqa: E402
                   random_ttn, tfim_tree_mpo, tree_mpo)

__all__ = [
    "Context", "DeviceTensor", "HeffMap", "env_update", "factorize", "krylov_expm", "lanczos_eig", "qr", "theta",
    "heisenberg_mpo", "product_mps", "random_mps", "xxz_mpo", "ProjTTN", "TTN", "TTNO", "expect_mpo",
    "orthogonalize", "expect_local", "Graph", "euler_sweep", "euler_tour_edges", "euler_tour_vertices", "path_graph",
    "post_order_dfs_plan", "tdvp_regions", "truncation_parameters", "EigsolveProblem", "dmrg", "eigsolve",
    "eigsolve_solver", "extracter", "inserter", "optimal_map", "region_iterator", "sweep_iterator",
    "sweep_solve", "updater", "ApplyExpProblem", "applyexp", "current_time", "exponentiate_solver",
    "process_real_times", "runge_kutta_solver", "tdvp", "NetMap", "permute", "tensordot", "TreeGraph",
    "TreeProjTTN", "TreeTTN", "TreeTTNO", "heisenberg_tree_mpo", "product_ttn", "random_ttn", "tfim_tree_mpo",
    "tree_mpo", "QNIndex", "QNMPO", "QNMPS", "QNProjTTN", "QNTensor", "qn_product_mps", "spin_site_index",
]
