"""Region plans of the oracle vs the reference's own Euler-tour test
(test/test_euler_tour.jl:7-26) and hand-derived chain orders."""
import pytest

from oracle import plans


def build_tree(nbranch=3, nbranch_sites=3):
    verts = [(0, 0)] + [(b, s) for b in range(1, nbranch + 1) for s in range(1, nbranch_sites + 1)]
    edges = []
    for b in range(1, nbranch + 1):
        edges.append(((0, 0), (b, 1)))
        for s in range(2, nbranch_sites + 1):
            edges.append(((b, s - 1), (b, s)))
    return verts, edges


def test_euler_tour_reference_invariants():
    verts, edges = build_tree()
    tour = plans.euler_tour_edges(verts, edges, (0, 0))
    for j in range(len(tour) - 1):
        assert tour[j][1] == tour[j + 1][0]
    for (a, b) in edges:
        assert (a, b) in tour and (b, a) in tour
    vt = plans.euler_tour_vertices(verts, edges, (0, 0))
    assert set(vt) == set(verts)
    assert len(tour) == 2 * len(edges)


def test_chain_euler_from_last_site():
    assert plans.euler_sweep_chain(4, 3) == [(3, 2), (2, 1), (1, 0), (0, 1), (1, 2), (2, 3)]
    assert plans.euler_sweep_chain(4, 0) == [(0, 1), (1, 2), (2, 3), (3, 2), (2, 1), (1, 0)]


def test_get_or_last():
    assert [plans.get_or_last([10, 20, 40], i) for i in range(1, 6)] == [10, 20, 40, 40, 40]
    assert plans.get_or_last(1e-10, 7) == 1e-10


def test_tdvp_plan_structure():
    vs, es = list(range(4)), [(0, 1), (1, 2), (2, 3)]
    p = plans.first_order_sweep(vs, es, 3, 0.1, False, 2)
    assert [r for r, _ in p] == [[0, 1], [1], [1, 2], [2], [2, 3]]
    assert [t for _, t in p] == [0.1, -0.1, 0.1, -0.1, 0.1]
    rev = plans.first_order_sweep(vs, es, 3, 0.1, True, 2)
    assert [r for r, _ in rev] == [[3, 2], [2], [2, 1], [1], [1, 0]]
    full = plans.tdvp_regions(vs, es, 3, 0.05, 4, 2)
    assert len(full) == 6 * 5
    assert sum(t for r, t in full if len(r) == 2) / 3 == pytest.approx(0.05)
    with pytest.raises(ValueError):
        plans.tdvp_sub_time_steps(3)
