"""Host-side serving logic (no device calls)."""


def test_footprint_grows_with_size_and_bond_dimension():
    from networksolvers_amd.serving import SolveRequest, footprint_bytes
    a = footprint_bytes(SolveRequest(0, 16, 1.0, 64))
    b = footprint_bytes(SolveRequest(1, 128, 1.0, 64))
    c = footprint_bytes(SolveRequest(2, 128, 1.0, 256))
    assert a < b < c
    assert c < 4 << 30


def test_stream_is_seeded_and_in_range():
    from networksolvers_amd.serving import make_stream
    s1 = make_stream(200, rate=5.0, seed=7)
    s2 = make_stream(200, rate=5.0, seed=7)
    assert [(r.N, r.delta, r.maxdim, r.seed, r.arrival) for r in s1] == \
           [(r.N, r.delta, r.maxdim, r.seed, r.arrival) for r in s2]
    assert all(16 <= r.N <= 128 and 0 <= r.delta <= 2 and r.maxdim in (64, 128, 256) for r in s1)
    assert all(b.arrival >= a.arrival for a, b in zip(s1, s1[1:]))
    mean_gap = s1[-1].arrival / len(s1)
    assert 0.1 < mean_gap < 0.4


def test_latency_stats():
    from networksolvers_amd.serving import SolveResult, latency_stats
    res = [SolveResult(i, -1.0, -1.0, 4, 0.0, 0.0, float(i + 1)) for i in range(100)]
    st = latency_stats(res)
    assert abs(st["p50_s"] - 50.5) < 1e-9 and 99 <= st["p99_s"] <= 100
