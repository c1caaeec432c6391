"""Region plans and truncation schedules (host integer logic).

Mirrors truncation_parameters.jl, region_plans/euler_tour.jl, euler_plans.jl,
dfs_plans.jl and tdvp_region_plans.jl.  A graph is a ``Graph`` (vertices in
insertion order, edges as ordered pairs); neighbour order follows the edge list
as in ``compute_adjacencies`` (euler_tour.jl:4-11).
"""
from __future__ import annotations

from dataclasses import dataclass, field

# ------------------------------------------------------------------ truncation
DEFAULT_MAXDIM = None  # typemax(Int)
DEFAULT_MINDIM = 1
DEFAULT_CUTOFF = 0.0


def get_or_last(x, i):
    """truncation_parameters.jl:5: x[i] (1-based) for vector schedules, last(x) once
    i >= length(x); scalars are their own last element."""
    if isinstance(x, (list, tuple)):
        return x[-1] if i >= len(x) else x[i - 1]
    return x


def truncation_parameters(sweep, *, cutoff=DEFAULT_CUTOFF, maxdim=DEFAULT_MAXDIM, mindim=DEFAULT_MINDIM):
    """truncation_parameters.jl:7-14 -> dict(cutoff, mindim, maxdim) for this sweep."""
    return {"cutoff": get_or_last(cutoff, sweep), "mindim": get_or_last(mindim, sweep),
            "maxdim": get_or_last(maxdim, sweep)}


# ---------------------------------------------------------------------- graphs
@dataclass
class Graph:
    vertices: list
    edges: list = field(default_factory=list)

    def neighbors(self, v):
        return compute_adjacencies(self)[v]

    def is_leaf(self, v):
        return len(self.neighbors(v)) == 1

    def is_path(self):
        """True when the graph is the chain 0-1-...-(N-1) in vertex order."""
        n = len(self.vertices)
        return self.vertices == list(range(n)) and sorted(tuple(sorted(e)) for e in self.edges) == \
            [(i, i + 1) for i in range(n - 1)]


def path_graph(N: int) -> Graph:
    return Graph(list(range(N)), [(i, i + 1) for i in range(N - 1)])


def compute_adjacencies(G: Graph):
    """euler_tour.jl:4-11."""
    adj = {v: [] for v in G.vertices}
    for (a, b) in G.edges:
        adj[a].append(b)
        adj[b].append(a)
    return adj


def default_root_vertex(G: Graph):
    """NamedGraphs ``default_root_vertex``: the last leaf in vertex order (on a
    chain, the last site; the recalled rule of the dependency)."""
    adj = compute_adjacencies(G)
    leaves = [v for v in G.vertices if len(adj[v]) <= 1]
    return leaves[-1] if leaves else G.vertices[-1]


# ------------------------------------------------------------------ Euler tour
def euler_tour_edges(G: Graph, start_vertex):
    """euler_tour.jl:13-42 -- DFS emitting every tree edge once in each direction."""
    adj = compute_adjacencies(G)
    visited = set()
    tour = []
    stack = [start_vertex]
    while stack:
        u = stack[-1]
        pushed = False
        for v in adj[u]:
            if (u, v) not in visited:
                visited.add((u, v))
                visited.add((v, u))
                tour.append((u, v))
                stack.append(v)
                pushed = True
                break
        if not pushed:
            stack.pop()
            if stack:
                tour.append((u, stack[-1]))
    return tour


def euler_tour_vertices(G: Graph, start_vertex):
    """euler_tour.jl:44-48."""
    edges = euler_tour_edges(G, start_vertex)
    if not edges:
        return []
    return [edges[0][0]] + [e[1] for e in edges]


def euler_sweep(G: Graph, *, nsites, root_vertex=None, **sweep_kwargs):
    """euler_plans.jl:4-13 -> [(region, sweep_kwargs)]."""
    root = default_root_vertex(G) if root_vertex is None else root_vertex
    if nsites == 1:
        return [([v], dict(sweep_kwargs)) for v in euler_tour_vertices(G, root)]
    if nsites == 2:
        return [([a, b], dict(sweep_kwargs)) for (a, b) in euler_tour_edges(G, root)]
    raise ValueError(f"euler_sweep: nsites={nsites} is not supported (1 or 2)")


# --------------------------------------------------------------- post-order DFS
def post_order_dfs_vertices(G: Graph, root):
    """NamedGraphs ``post_order_dfs_vertices``: each vertex after its subtree,
    children in adjacency order."""
    adj = compute_adjacencies(G)
    out, parent = [], {root: None}
    stack = [(root, iter(adj[root]))]
    while stack:
        v, it = stack[-1]
        nxt = None
        for c in it:
            if c not in parent:
                nxt = c
                break
        if nxt is None:
            out.append(v)
            stack.pop()
        else:
            parent[nxt] = v
            stack.append((nxt, iter(adj[nxt])))
    return out, parent


def post_order_dfs_edges(G: Graph, root):
    """NamedGraphs ``post_order_dfs_edges``: (v -> parent(v)) in post order."""
    order, parent = post_order_dfs_vertices(G, root)
    return [(v, parent[v]) for v in order if parent[v] is not None]


def post_order_dfs_plan(G: Graph, *, nsites, root_vertex=None, **sweep_kwargs):
    """dfs_plans.jl:5-16."""
    root = default_root_vertex(G) if root_vertex is None else root_vertex
    if nsites == 1:
        order, _ = post_order_dfs_vertices(G, root)
        return [([v], dict(sweep_kwargs)) for v in order]
    if nsites == 2:
        return [([a, b], dict(sweep_kwargs)) for (a, b) in post_order_dfs_edges(G, root)]
    raise ValueError(f"post_order_dfs_plan: nsites={nsites} is not supported")


def post_order_dfs_sweep(G: Graph, **kws):
    """dfs_plans.jl:18-22."""
    fwd = post_order_dfs_plan(G, **kws)
    rev = [(list(reversed(r)), k) for (r, k) in reversed(fwd)]
    return fwd + rev


# ------------------------------------------------------------------------ TDVP
def tdvp_sub_time_steps(tdvp_order):
    """tdvp_region_plans.jl:1-12."""
    if tdvp_order == 1:
        return [1.0]
    if tdvp_order == 2:
        return [0.5, 0.5]
    if tdvp_order == 4:
        s = (2 - 2 ** (1 / 3)) ** (-1)
        return [s / 2, s / 2, 0.5 - s, 0.5 - s, s / 2, s / 2]
    raise ValueError(f"TDVP order of {tdvp_order} not supported")


def first_order_sweep(G: Graph, time_step, reverse=False, *, updater_kwargs, nsites, **kws):
    """tdvp_region_plans.jl:14-32: forward post-order sweep with a backward
    1-site step (-time_step) after every 2-site region except the last."""
    basic = post_order_dfs_plan(G, nsites=nsites, **kws)
    uk = {"nsites": nsites, "time_step": time_step, **updater_kwargs}
    sweep = []
    for j, (region, region_kws) in enumerate(basic):
        sweep.append((region, {"nsites": nsites, "updater_kwargs": uk, **region_kws}))
        if len(region) == 2 and j < len(basic) - 1:
            rev_kwargs = {**uk, "time_step": -uk["time_step"]}
            sweep.append(([region[-1]], {"updater_kwargs": rev_kwargs, **region_kws}))
    if reverse:
        sweep = [(list(reversed(r)), k) for (r, k) in reversed(sweep)]
    return sweep


def tdvp_regions(G: Graph, time_step, *, updater_kwargs, tdvp_order, nsites, **kws):
    """tdvp_region_plans.jl:34-44."""
    plan = []
    for step, weight in enumerate(tdvp_sub_time_steps(tdvp_order), start=1):
        plan += first_order_sweep(G, weight * time_step, step % 2 == 0, updater_kwargs=updater_kwargs,
                                  nsites=nsites, **kws)
    return plan
