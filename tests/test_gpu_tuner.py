"""Tile autotuner of the block-program GEMMs (nsmi_tune_mode / nsmi_tune_table).

The grouped GEMM runs with 32 x 32 or 64 x 64 output tiles per workgroup.  Every
output element sums its segments and K in the same order under either tile, so the
choice must not change a bit of the result: forced tiles, the shape heuristic and the
autotuner (which alternates the two while it times them) give identical arrays, equal
to the dense contraction of the oracle.  The tuner has no counterpart in the
reference (ITensors contracts each block pair with BLAS)."""
import numpy as np
import pytest

from oracle import chain, models

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import networksolvers_amd as ns
    c = ns.Context(0)
    yield c
    c.close()


def _rand_qn(ctx, indices, dirs, seed):
    from networksolvers_amd import qn
    t = qn.QNTensor(ctx, indices, dirs, 0)
    t.data = ctx.upload(np.random.default_rng(seed).standard_normal(max(t.size, 1)))
    return t


def _problem(ctx):
    """A 2-site H_eff whose charge blocks reach 70 x 70 (several tiles of either size)."""
    from networksolvers_amd import qn
    site = qn.spin_site_index(2)
    l = qn.QNIndex([(-2, 40), (0, 70), (2, 36)])
    r = qn.QNIndex([(-4, 9), (-2, 44), (0, 66), (2, 41), (4, 12)])
    H = qn.QNMPO(ctx, models.heisenberg_mpo(4), site)
    L = _rand_qn(ctx, [l, H.links[1], l], [1, -1, -1], 11)
    R = _rand_qn(ctx, [r, H.links[3], r], [1, -1, -1], 12)
    th = _rand_qn(ctx, [l, site, site, r], [1, 1, 1, -1], 13)
    return H, L, R, th


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_tile_choice_is_bit_identical(ctx, precision):
    from networksolvers_amd import qn
    H, L, R, th = _problem(ctx)
    op = qn.heff2(L, H.W[1], H.W[2], R, th, H.c[1], H.c[3], ctx, precision)
    outs = {}
    try:
        for mode in ("tile32", "tile64", "heuristic"):
            ctx.tune_mode(mode)
            outs[mode] = qn.QNTensor(ctx, th.indices, th.dirs, 0, th.keys, op(th.data)).to_dense()
        ctx.tune_mode("auto")
        # enough applies for every class of the program to be timed with both tiles
        autos = [qn.QNTensor(ctx, th.indices, th.dirs, 0, th.keys, op(th.data)).to_dense() for _ in range(10)]
    finally:
        ctx.tune_mode("auto")
        op.close()
    ref = outs["tile32"]
    for mode, y in outs.items():
        assert np.array_equal(y, ref), mode
    for i, y in enumerate(autos):
        assert np.array_equal(y, ref), f"autotuned apply {i}"
    want = chain.heff2_apply(L.to_dense(), H.W[1], H.W[2], R.to_dense(), th.to_dense())
    err = np.linalg.norm(ref - want) / np.linalg.norm(want)
    assert err < (1e-12 if precision == "fp64" else 2e-5)


def test_tuner_decides_each_class(ctx):
    """After repeated runs every class with both tiles has three timed samples of each
    and a decision; the table reports launches, samples and best times consistently."""
    from networksolvers_amd import qn
    H, L, R, th = _problem(ctx)
    ctx.tune_mode("auto")
    op = qn.heff2(L, H.W[1], H.W[2], R, th, H.c[1], H.c[3], ctx)
    try:
        for _ in range(10):
            op(th.data)
        rows = ctx.tune_table()
    finally:
        op.close()
    assert rows, "no GEMM class was tuned"
    decided = [r for r in rows if r["choice"] != -1]
    assert decided
    for r in decided:
        assert r["choice"] in (32, 64)
        assert min(r["samples"]) >= 3 and min(r["issued"]) >= 3
        assert min(r["best_ns"]) > 0
        faster = 64 if r["best_ns"][1] < r["best_ns"][0] else 32
        assert r["choice"] == faster or r["best_ns"][0] == r["best_ns"][1]
    for r in rows:
        assert r["samples"][0] <= r["issued"][0] and r["samples"][1] <= r["issued"][1]


def test_tune_mode_rejects_unknown(ctx):
    from networksolvers_amd._lib import NsmiError
    with pytest.raises(NsmiError, match="unknown mode"):
        ctx.check(ctx.lib.nsmi_tune_mode(ctx.h, 7))
    ctx.tune_mode("auto")
