"""Region plans restated from NetworkSolvers.jl (test infrastructure only).

Graphs are given as (vertices, edges) with vertices in insertion order and edges
as ordered pairs; neighbour order follows the edge list (euler_tour.jl:4-11).
"""
from __future__ import annotations


def get_or_last(x, i):
    """truncation_parameters.jl:5 -- x[i] for a vector schedule (1-based sweep i),
    its last entry once i >= length(x); scalars pass through."""
    if isinstance(x, (list, tuple)):
        return x[-1] if i >= len(x) else x[i - 1]
    return x


def compute_adjacencies(vertices, edges):
    """euler_tour.jl:4-11."""
    adj = {v: [] for v in vertices}
    for (a, b) in edges:
        adj[a].append(b)
        adj[b].append(a)
    return adj


def euler_tour_edges(vertices, edges, start_vertex):
    """euler_tour.jl:13-42 -- iterative DFS emitting each tree edge in both directions."""
    adj = compute_adjacencies(vertices, edges)
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


def euler_tour_vertices(vertices, edges, start_vertex):
    """euler_tour.jl:44-48."""
    tour = euler_tour_edges(vertices, edges, start_vertex)
    if not tour:
        return []
    return [tour[0][0]] + [e[1] for e in tour]


def chain_graph(N):
    return list(range(N)), [(i, i + 1) for i in range(N - 1)]


def euler_sweep_chain(N, root_vertex):
    """euler_plans.jl:4-13 with nsites=2 on a chain: [src(e), dst(e)] per tour edge."""
    vs, es = chain_graph(N)
    return euler_tour_edges(vs, es, root_vertex)


def post_order_dfs_vertices(vertices, edges, root):
    """NamedGraphs ``post_order_dfs_vertices`` (used by dfs_plans.jl:9): children are
    visited in adjacency order, each vertex emitted after its subtree.  The
    neighbour order of the dependency is a recalled detail (parity unpinned)."""
    adj = compute_adjacencies(vertices, edges)
    out = []
    parent = {root: None}
    stack = [(root, iter(adj[root]))]
    while stack:
        v, it = stack[-1]
        nxt = None
        for c in it:
            if c != parent[v] and c not in parent:
                nxt = c
                break
        if nxt is None:
            out.append(v)
            stack.pop()
        else:
            parent[nxt] = v
            stack.append((nxt, iter(adj[nxt])))
    return out, parent


def post_order_dfs_edges(vertices, edges, root):
    """NamedGraphs ``post_order_dfs_edges`` (dfs_plans.jl:12): (v -> parent(v)) for
    every non-root vertex in post order."""
    order, parent = post_order_dfs_vertices(vertices, edges, root)
    return [(v, parent[v]) for v in order if parent[v] is not None]


def tdvp_sub_time_steps(order):
    """tdvp_region_plans.jl:1-12."""
    if order == 1:
        return [1.0]
    if order == 2:
        return [0.5, 0.5]
    if order == 4:
        s = (2 - 2 ** (1 / 3)) ** (-1)
        return [s / 2, s / 2, 0.5 - s, 0.5 - s, s / 2, s / 2]
    raise ValueError(f"TDVP order of {order} not supported")


def first_order_sweep(vertices, edges, root, time_step, reverse, nsites):
    """tdvp_region_plans.jl:14-32 -> list of (region, time_step)."""
    if nsites == 1:
        order, _ = post_order_dfs_vertices(vertices, edges, root)
        fwd = [[v] for v in order]
    else:
        fwd = [[a, b] for (a, b) in post_order_dfs_edges(vertices, edges, root)]
    sweep = []
    for j, region in enumerate(fwd):
        sweep.append((region, time_step))
        if len(region) == 2 and j < len(fwd) - 1:
            sweep.append(([region[-1]], -time_step))
    if reverse:
        sweep = [(list(reversed(r)), t) for (r, t) in reversed(sweep)]
    return sweep


def tdvp_regions(vertices, edges, root, time_step, order, nsites):
    """tdvp_region_plans.jl:34-44."""
    plan = []
    for step, w in enumerate(tdvp_sub_time_steps(order), start=1):
        plan += first_order_sweep(vertices, edges, root, w * time_step, step % 2 == 0, nsites)
    return plan
