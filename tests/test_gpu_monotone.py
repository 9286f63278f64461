"""Monotone constraints on the GPU (tree.monotone_constraints): trees grown with
directions equal the oracle's bit for bit (the finder's candidate skip,
src/split.cpp:143-153; the boxes of propagate_monotone_bounds,
src/grower_impl.hpp:47-75; leaf values clamped to them, include/bonsai/split.hpp:141-150),
every trained tree is monotone in the constrained features, and the levelwise grower
refuses constraints (src/grower_impl.hpp:583-588).
"""
import numpy as np
import pytest

from oracle_lib import HIST_F32_FIXED, HIST_INT, cfg as orc_cfg
from test_gpu_depthwise import assert_same_tree, ctx_for, grad_of, tall_data
from test_gpu_growers import LW_F32_CHUNK
from test_oracle_monotone import monotone_violations

pytestmark = pytest.mark.gpu

DIRS = np.zeros(28, np.int8)
DIRS[[0, 1, 10]] = [1, -1, 1]
DIRS[24] = -1

KW = {"hip_depthwise": dict(max_depth=7), "hip_leafwise": dict(max_depth=10, max_leaves=63)}


@pytest.fixture(scope="module")
def tall():
    return tall_data(40000, seed_cfg=6)


def oracle_tree(orc, grower, bins, nb, g, h, gq, hq, sc, quantized, c):
    orc.set_monotone(DIRS)
    try:
        if grower == "hip_depthwise":
            if quantized:
                return orc.grow_depthwise(bins, nb, gq=gq, hq=hq, scales=sc, c=c, mode=HIST_INT)
            return orc.grow_depthwise(bins, nb, g=g, h=h, c=c, mode=HIST_F32_FIXED)
        if quantized:
            return orc.grow_leafwise(bins, nb, gq=gq, hq=hq, scales=sc, c=c, mode=HIST_INT)
        return orc.grow_leafwise(bins, nb, g=g, h=h, c=c, mode=HIST_F32_FIXED, chunk=LW_F32_CHUNK)
    finally:
        orc.set_monotone(None)


@pytest.mark.parametrize("quantized", [True, False])
@pytest.mark.parametrize("grower", ["hip_depthwise", "hip_leafwise"])
def test_monotone_tree_matches_oracle(hip, orc, tall, grower, quantized):
    bins, nb, X, y = tall
    g, h = grad_of(orc, y, seed=23)
    ctx = ctx_for(hip, bins, nb, grower=grower, quantized=quantized, **KW[grower])
    gq = hq = sc = None
    if quantized:
        gq, hq, sc = ctx.set_gradients(g, h)
    else:
        ctx.set_gradients(g, h)
    free, _ = ctx.grow_tree()
    ctx.set_monotone(DIRS)
    nodes, lor = ctx.grow_tree()
    ctx.close()
    ref, rlor = oracle_tree(orc, grower, bins, nb, g, h, gq, hq, sc, quantized, orc_cfg(**KW[grower]))
    assert_same_tree(nodes, ref)
    assert np.array_equal(lor, rlor)
    assert nodes.tobytes() != free.tobytes()  # the constraints changed the tree
    for f in np.nonzero(DIRS)[0]:
        assert monotone_violations(nodes, bins, nb, f, int(DIRS[f]), n_rows=60) == 0


@pytest.mark.parametrize("grower", ["hip_depthwise", "hip_leafwise"])
def test_monotone_training_graph_equals_eager(hip, orc, tall, grower):
    bins, nb, X, y = tall
    res = []
    for use_graph in (True, False):
        ctx = ctx_for(hip, bins, nb, grower=grower, quantized=True, use_graph=use_graph, learning_rate=0.1,
                      **KW[grower])
        ctx.upload_labels(y)
        ctx.set_scores(np.full(y.size, orc.init_score(1, y), np.float32))
        ctx.set_monotone(DIRS)
        ctx.train(4)
        res.append((ctx.trees(), ctx.get_scores()))
        ctx.close()
    (ta, sa), (tb, sb) = res
    assert len(ta) == len(tb) == 4
    for a, b in zip(ta, tb):
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8))
        for f in np.nonzero(DIRS)[0]:
            assert monotone_violations(a, bins, nb, f, int(DIRS[f]), n_rows=30) == 0
    assert np.array_equal(sa.view(np.uint32), sb.view(np.uint32))


def test_monotone_refused_by_levelwise_and_clearable(hip, orc, tall):
    bins, nb, X, y = tall
    ctx = ctx_for(hip, bins, nb, grower="hip_levelwise", quantized=True, max_depth=4)
    with pytest.raises(hip.ConfigError):
        ctx.set_monotone(DIRS)
    ctx.set_monotone(np.zeros(28, np.int8))  # all free: accepted
    ctx.close()
    g, h = grad_of(orc, y, seed=29)
    out = []
    for dirs in (None, DIRS):
        ctx = ctx_for(hip, bins, nb, grower="hip_depthwise", quantized=True, max_depth=6)
        ctx.set_gradients(g, h)
        if dirs is not None:
            ctx.set_monotone(dirs)
            ctx.set_monotone(None)
        out.append(ctx.grow_tree()[0])
        ctx.close()
    assert out[0].tobytes() == out[1].tobytes()


def test_reconfigure_owns_the_constraints(hip, orc, tall):
    """configure clears the constraints an earlier fit or setter left: a depthwise
    context with monotone and interaction constraints reconfigures to levelwise and
    grows the unconstrained oblivious tree; back on depthwise the free tree."""
    bins, nb, X, y = tall
    g, h = grad_of(orc, y, seed=31)
    ctx = ctx_for(hip, bins, nb, grower="hip_depthwise", quantized=True, max_depth=5,
                  monotone_constraints=tuple(int(v) for v in DIRS), interaction_constraints=("0,1", "2+3"))
    ctx.set_gradients(g, h)
    constrained, _ = ctx.grow_tree()
    ctx.configure(hip.FitConfig(grower="hip_levelwise", quantized=True, max_depth=4))
    ctx.set_gradients(g, h)
    lv, _ = ctx.grow_tree()
    ref = hip.HipContext(0)
    ref.upload_bins(bins, nb)
    ref.configure(hip.FitConfig(grower="hip_levelwise", quantized=True, max_depth=4))
    ref.set_gradients(g, h)
    assert lv.tobytes() == ref.grow_tree()[0].tobytes()
    ref.configure(hip.FitConfig(grower="hip_depthwise", quantized=True, max_depth=5))
    ref.set_gradients(g, h)
    free, _ = ref.grow_tree()
    ref.close()
    ctx.configure(hip.FitConfig(grower="hip_depthwise", quantized=True, max_depth=5))
    ctx.set_gradients(g, h)
    again, _ = ctx.grow_tree()
    ctx.close()
    assert again.tobytes() == free.tobytes() != constrained.tobytes()


def test_refused_configure_keeps_the_previous_fit(hip, orc, tall):
    """Through the C ABI: with directions set, bh_configure refuses the levelwise grower
    before anything changes, so the constrained depthwise fit still grows its tree."""
    import ctypes as C
    bins, nb, X, y = tall
    g, h = grad_of(orc, y, seed=37)
    ctx = ctx_for(hip, bins, nb, grower="hip_depthwise", quantized=True, max_depth=6)
    ctx.set_monotone(DIRS)
    gq, hq, sc = ctx.set_gradients(g, h)
    before, _ = ctx.grow_tree()
    bad = hip.FitConfig(grower="hip_levelwise", quantized=True, max_depth=4).to_c()
    assert hip.lib().bh_configure(ctx._h, C.byref(bad)) > 0
    assert b"levelwise" in hip.lib().bh_last_error()
    ctx.set_gradients(g, h)
    after, _ = ctx.grow_tree()
    ctx.close()
    assert after.tobytes() == before.tobytes()
    ref, _ = oracle_tree(orc, "hip_depthwise", bins, nb, g, h, gq, hq, sc, True, orc_cfg(max_depth=6))
    assert_same_tree(after, ref)
