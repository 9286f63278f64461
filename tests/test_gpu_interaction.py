"""Interaction constraints on the GPU (tree.interaction_constraints): trees grown with
groups equal the oracle's bit for bit (allowed_features, src/grower_impl.hpp:111-137;
propagate_interaction_state, :143-161; the finder's skip, src/split.cpp:111-114), every
path stays inside a group, and the levelwise grower and out-of-range ids are refused.
"""
import numpy as np
import pytest

from oracle_lib import HIST_F32_FIXED, HIST_INT, cfg as orc_cfg
from test_gpu_depthwise import assert_same_tree, ctx_for, grad_of, tall_data
from test_gpu_growers import LW_F32_CHUNK
from test_oracle_interaction import violations

pytestmark = pytest.mark.gpu

GROUPS = [[0, 1, 2, 10], [3, 4, 24], [2, 5, 6, 7, 8, 9]]
KW = {"hip_depthwise": dict(max_depth=7), "hip_leafwise": dict(max_depth=10, max_leaves=63)}


@pytest.fixture(scope="module")
def tall():
    return tall_data(40000, seed_cfg=7)


def oracle_tree(orc, grower, bins, nb, g, h, gq, hq, sc, quantized, c):
    orc.set_interaction(GROUPS)
    try:
        fn = orc.grow_depthwise if grower == "hip_depthwise" else orc.grow_leafwise
        if quantized:
            return fn(bins, nb, gq=gq, hq=hq, scales=sc, c=c, mode=HIST_INT)
        if grower == "hip_leafwise":
            return fn(bins, nb, g=g, h=h, c=c, mode=HIST_F32_FIXED, chunk=LW_F32_CHUNK)
        return fn(bins, nb, g=g, h=h, c=c, mode=HIST_F32_FIXED)
    finally:
        orc.set_interaction(None)


@pytest.mark.parametrize("quantized", [True, False])
@pytest.mark.parametrize("grower", ["hip_depthwise", "hip_leafwise"])
def test_interaction_tree_matches_oracle(hip, orc, tall, grower, quantized):
    bins, nb, X, y = tall
    g, h = grad_of(orc, y, seed=31)
    ctx = ctx_for(hip, bins, nb, grower=grower, quantized=quantized, **KW[grower])
    gq = hq = sc = None
    if quantized:
        gq, hq, sc = ctx.set_gradients(g, h)
    else:
        ctx.set_gradients(g, h)
    free, _ = ctx.grow_tree()
    ctx.set_interaction_constraints(["0,1,2,10", "3+4+24", [9, 8, 7, 6, 5, 2]])
    nodes, lor = ctx.grow_tree()
    ctx.close()
    ref, rlor = oracle_tree(orc, grower, bins, nb, g, h, gq, hq, sc, quantized, orc_cfg(**KW[grower]))
    assert_same_tree(nodes, ref)
    assert np.array_equal(lor, rlor)
    assert violations(free, GROUPS) > 0
    assert violations(nodes, GROUPS) == 0


@pytest.mark.parametrize("grower", ["hip_depthwise", "hip_leafwise"])
def test_interaction_training_graph_equals_eager(hip, orc, tall, grower):
    bins, nb, X, y = tall
    res = []
    for use_graph in (True, False):
        ctx = ctx_for(hip, bins, nb, grower=grower, quantized=True, use_graph=use_graph, learning_rate=0.1,
                      **KW[grower])
        ctx.upload_labels(y)
        ctx.set_scores(np.full(y.size, orc.init_score(1, y), np.float32))
        ctx.set_interaction_constraints(GROUPS)
        ctx.set_monotone([1, 0, 0, -1])  # both kinds of constraint at once
        ctx.train(4)
        res.append((ctx.trees(), ctx.get_scores()))
        ctx.close()
    (ta, sa), (tb, sb) = res
    assert len(ta) == len(tb) == 4
    for a, b in zip(ta, tb):
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8))
        assert violations(a, GROUPS) == 0
    assert np.array_equal(sa.view(np.uint32), sb.view(np.uint32))


def test_interaction_refusals_and_clearing(hip, orc, tall):
    bins, nb, X, y = tall
    ctx = ctx_for(hip, bins, nb, grower="hip_levelwise", quantized=True, max_depth=4)
    with pytest.raises(hip.ConfigError):
        ctx.set_interaction_constraints(GROUPS)
    ctx.close()
    ctx = ctx_for(hip, bins, nb, grower="hip_depthwise", quantized=True, max_depth=5)
    with pytest.raises(hip.ConfigError):
        ctx.set_interaction_constraints([[0, 28]])
    g, h = grad_of(orc, y, seed=37)
    ctx.set_gradients(g, h)
    a = ctx.grow_tree()[0]
    ctx.set_interaction_constraints(GROUPS)
    ctx.set_interaction_constraints(None)
    b = ctx.grow_tree()[0]
    ctx.close()
    assert a.tobytes() == b.tobytes()


def test_fit_config_fields_equal_the_setters(hip, orc, tall):
    """FitConfig's TreeConfig fields (feature_fraction / feature_seed,
    monotone_constraints, interaction_constraints) reach the device through configure."""
    bins, nb, X, y = tall
    g, h = grad_of(orc, y, seed=41)
    fields = dict(feature_fraction=0.6, feature_seed=9, monotone_constraints=(1, 0, -1),
                  interaction_constraints=("0,1,2,10", "3+4+24", "2,5,6,7,8,9"))
    a_ctx = ctx_for(hip, bins, nb, grower="hip_depthwise", quantized=True, max_depth=6, **fields)
    a_ctx.set_gradients(g, h)
    a = a_ctx.grow_tree()[0]
    a_ctx.close()
    b_ctx = ctx_for(hip, bins, nb, grower="hip_depthwise", quantized=True, max_depth=6)
    b_ctx.set_gradients(g, h)
    b_ctx.set_feature_fraction(0.6, seed=9)
    b_ctx.set_monotone([1, 0, -1])
    b_ctx.set_interaction_constraints(GROUPS)
    b = b_ctx.grow_tree()[0]
    b_ctx.close()
    assert a.tobytes() == b.tobytes()
    with pytest.raises(hip.ConfigError):
        ctx_for(hip, bins, nb, grower="hip_levelwise", quantized=True, max_depth=4, monotone_constraints=(1,))
