"""GPU parity of the leafwise and levelwise growers against the CPU oracle.

Quantized trees are bit-exact (same quantized gradients); FP32 trees are bit-exact
against the oracle restating the device's fixed summation order.
"""
import numpy as np
import pytest

from oracle_lib import CHUNK_ADAPTIVE, HIST_F32_FIXED, HIST_INT, cfg as orc_cfg
from test_gpu_depthwise import assert_same_tree, ctx_for, grad_of, tall_data

pytestmark = pytest.mark.gpu

# Leafwise FP32 histograms sum over leaf-sized units (at most 256 per built node at these
# shapes); the oracle follows the same rule.
LW_F32_CHUNK = CHUNK_ADAPTIVE | 256


@pytest.fixture(scope="module")
def tall():
    return tall_data(40000, seed_cfg=2)


def wide_data(n=6000, f=300):
    from bonsai_amd import synth
    from oracle_lib import load_oracle
    orc = load_oracle()
    X, y = synth.wide_bins(n, f)
    cuts = [orc.fit_cuts(X[:, j], 255) for j in range(f)]
    bins = np.stack([orc.transform(cuts[j], X[:, j]) for j in range(f)], axis=1)
    nb = np.array([c.size for c in cuts], np.uint16)
    return bins, nb, X, y


@pytest.fixture(scope="module")
def wide():
    return wide_data()


@pytest.mark.parametrize("leaves,depth", [(2, 12), (31, 12), (63, 5), (255, 12)])
def test_grow_leafwise_int_bit_exact(hip, orc, tall, leaves, depth):
    bins, nb, X, y = tall
    g, h = grad_of(orc, y, seed=leaves)
    ctx = ctx_for(hip, bins, nb, grower="hip_leafwise", quantized=True, max_depth=depth, max_leaves=leaves)
    gq, hq, sc = ctx.set_gradients(g, h)
    nodes, lor = ctx.grow_tree()
    ref, rlor = orc.grow_leafwise(bins, nb, gq=gq, hq=hq, scales=sc,
                                  c=orc_cfg(max_depth=depth, max_leaves=leaves), mode=HIST_INT)
    assert_same_tree(nodes, ref)
    assert np.array_equal(lor, rlor)


@pytest.mark.parametrize("leaves,depth", [(31, 12), (127, 7)])
def test_grow_leafwise_f32_matches_device_order_oracle(hip, orc, tall, leaves, depth):
    bins, nb, X, y = tall
    g, h = grad_of(orc, y, seed=100 + leaves)
    ctx = ctx_for(hip, bins, nb, grower="hip_leafwise", quantized=False, max_depth=depth, max_leaves=leaves)
    ctx.set_gradients(g, h)
    nodes, lor = ctx.grow_tree()
    ref, rlor = orc.grow_leafwise(bins, nb, g=g, h=h, c=orc_cfg(max_depth=depth, max_leaves=leaves),
                                  mode=HIST_F32_FIXED, chunk=LW_F32_CHUNK)
    assert_same_tree(nodes, ref)
    assert np.array_equal(lor, rlor)


@pytest.fixture(scope="module")
def tall600k():
    return tall_data(600000, seed_cfg=3)


# Large enough that the leaf-sized work units leave their minimum sizes: the root's
# partition chunks are 768 rows, the first children's 512, and a quantized child of
# more than 262,144 rows gets histogram units above 1024 rows, with many partials.
@pytest.mark.parametrize("quantized", [True, False])
@pytest.mark.parametrize("leaves,depth", [(15, 10), (63, 12)])
def test_grow_leafwise_600k_rows(hip, orc, tall600k, quantized, leaves, depth):
    bins, nb, X, y = tall600k
    g, h = grad_of(orc, y, seed=7 + leaves)
    ctx = ctx_for(hip, bins, nb, grower="hip_leafwise", quantized=quantized, max_depth=depth, max_leaves=leaves)
    c = orc_cfg(max_depth=depth, max_leaves=leaves)
    if quantized:
        gq, hq, sc = ctx.set_gradients(g, h)
        ref, rlor = orc.grow_leafwise(bins, nb, gq=gq, hq=hq, scales=sc, c=c, mode=HIST_INT)
    else:
        ctx.set_gradients(g, h)
        ref, rlor = orc.grow_leafwise(bins, nb, g=g, h=h, c=c, mode=HIST_F32_FIXED, chunk=LW_F32_CHUNK)
    nodes, lor = ctx.grow_tree()
    ctx.close()
    assert_same_tree(nodes, ref)
    assert np.array_equal(lor, rlor)


def levelwise_view(nodes, D):
    """Perfect dense tree -> (level splits, leaf values, leaf covers)."""
    levels = np.array([[nodes[(1 << k) - 1]["feature"], nodes[(1 << k) - 1]["bin"],
                        nodes[(1 << k) - 1]["default_left"]] for k in range(D)], np.int64).reshape(D, 3)
    leaves = nodes[(1 << D) - 1:]
    return levels, leaves["value"], leaves["cover"]


def check_levelwise(hip, orc, bins, nb, g, h, depth, quantized):
    ctx = ctx_for(hip, bins, nb, grower="hip_levelwise", quantized=quantized, max_depth=depth)
    if quantized:
        gq, hq, sc = ctx.set_gradients(g, h)
    else:
        ctx.set_gradients(g, h)
    nodes, lor = ctx.grow_tree()
    if quantized:
        levels, lv, lc, gains, rlor = orc.grow_levelwise(bins, nb, gq=gq, hq=hq, scales=sc,
                                                         c=orc_cfg(max_depth=depth), mode=HIST_INT)
    else:
        levels, lv, lc, gains, rlor = orc.grow_levelwise(bins, nb, g=g, h=h, c=orc_cfg(max_depth=depth),
                                                         mode=HIST_F32_FIXED)
    D = levels.shape[0]
    assert len(nodes) == (2 << D) - 1
    dl, dv, dc = levelwise_view(nodes, D)
    assert np.array_equal(dl, levels.astype(np.int64))
    assert np.array_equal(dv.view(np.uint32), lv.view(np.uint32))
    assert np.array_equal(dc, lc)
    assert np.array_equal(lor - ((1 << D) - 1), rlor)


@pytest.mark.parametrize("depth", [1, 4, 8])
def test_grow_levelwise_int_bit_exact(hip, orc, tall, depth):
    bins, nb, X, y = tall
    g, h = grad_of(orc, y, seed=200 + depth)
    check_levelwise(hip, orc, bins, nb, g, h, depth, True)


@pytest.mark.parametrize("depth", [3, 6])
def test_grow_levelwise_f32_matches_device_order_oracle(hip, orc, tall, depth):
    bins, nb, X, y = tall
    g, h = grad_of(orc, y, seed=300 + depth)
    check_levelwise(hip, orc, bins, nb, g, h, depth, False)


@pytest.mark.parametrize("quantized", [True, False])
def test_grow_levelwise_wide_rows(hip, orc, wide, quantized):
    """300 features: rows wider than the register-resident scatter, three feature slices."""
    bins, nb, X, y = wide
    rng = np.random.default_rng(3)
    g, h = orc.mse_grad(rng.normal(0, 1, y.size).astype(np.float32), y)
    check_levelwise(hip, orc, bins, nb, g, h, 6, quantized)


@pytest.mark.parametrize("quantized", [True, False])
def test_grow_depthwise_wide_rows(hip, orc, wide, quantized):
    bins, nb, X, y = wide
    rng = np.random.default_rng(4)
    g, h = orc.mse_grad(rng.normal(0, 1, y.size).astype(np.float32), y)
    ctx = ctx_for(hip, bins, nb, quantized=quantized, max_depth=5)
    if quantized:
        gq, hq, sc = ctx.set_gradients(g, h)
        ref, rlor = orc.grow_depthwise(bins, nb, gq=gq, hq=hq, scales=sc, c=orc_cfg(max_depth=5), mode=HIST_INT)
    else:
        ctx.set_gradients(g, h)
        ref, rlor = orc.grow_depthwise(bins, nb, g=g, h=h, c=orc_cfg(max_depth=5), mode=HIST_F32_FIXED)
    nodes, lor = ctx.grow_tree()
    assert_same_tree(nodes, ref)
    assert np.array_equal(lor, rlor)


@pytest.mark.parametrize("pair", [1, 0])
def test_grow_depthwise_int_50_features(hip, orc, pair):
    """52-byte rows: paired levels with 13 dwords per row in the histogram lanes."""
    bins, nb, X, y = wide_data(n=30000, f=50)
    rng = np.random.default_rng(5)
    g, h = orc.mse_grad(rng.normal(0, 1, y.size).astype(np.float32), y)
    ctx = ctx_for(hip, bins, nb, quantized=True, max_depth=6, min_data_in_leaf=5)
    ctx.set_tuning(pair_levels=pair)
    gq, hq, sc = ctx.set_gradients(g, h)
    ref, rlor = orc.grow_depthwise(bins, nb, gq=gq, hq=hq, scales=sc, c=orc_cfg(max_depth=6, min_data=5),
                                   mode=HIST_INT)
    nodes, lor = ctx.grow_tree()
    assert_same_tree(nodes, ref)
    assert np.array_equal(lor, rlor)


@pytest.mark.parametrize("grower", ["hip_leafwise", "hip_levelwise"])
def test_train_graph_equals_eager(hip, orc, tall, grower):
    bins, nb, X, y = tall
    res = []
    for use_graph in (True, False):
        ctx = ctx_for(hip, bins, nb, grower=grower, quantized=True, max_depth=6, max_leaves=31,
                      use_graph=use_graph, learning_rate=0.1)
        ctx.upload_labels(y)
        ctx.set_scores(np.full(y.size, orc.init_score(1, y), np.float32))
        ctx.train(3)
        ctx.train(2)
        res.append((ctx.trees(), ctx.get_scores()))
        ctx.close()
    (ta, sa), (tb, sb) = res
    assert len(ta) == len(tb) == 5
    for a, b in zip(ta, tb):
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8))
    assert np.array_equal(sa.view(np.uint32), sb.view(np.uint32))


@pytest.mark.parametrize("quantized", [True, False])
def test_grow_leafwise_wide_rows(hip, orc, quantized):
    """1,500-byte rows: partition chunks are capped at 128 rows by bytes, below the
    leaf-sized rule's 256-row minimum."""
    bins, nb, X, y = wide_data(n=4000, f=1500)
    rng = np.random.default_rng(6)
    g, h = orc.mse_grad(rng.normal(0, 1, y.size).astype(np.float32), y)
    ctx = ctx_for(hip, bins, nb, grower="hip_leafwise", quantized=quantized, max_depth=8, max_leaves=20,
                  min_data_in_leaf=5)
    c = orc_cfg(max_depth=8, max_leaves=20, min_data=5)
    if quantized:
        gq, hq, sc = ctx.set_gradients(g, h)
        ref, rlor = orc.grow_leafwise(bins, nb, gq=gq, hq=hq, scales=sc, c=c, mode=HIST_INT)
    else:
        ctx.set_gradients(g, h)
        ref, rlor = orc.grow_leafwise(bins, nb, g=g, h=h, c=c, mode=HIST_F32_FIXED, chunk=LW_F32_CHUNK)
    nodes, lor = ctx.grow_tree()
    ctx.close()
    assert_same_tree(nodes, ref)
    assert np.array_equal(lor, rlor)
