"""Sampled roots on the GPU: a tree grown on a row subset (the row_indices a booster's
sampler hands to grow, include/bonsai/booster.hpp:666-676; the root over them,
src/step/level.hpp:171-187) equals the oracle's tree grown on exactly those rows, and
every row, sampled or not, gets its leaf from the finished tree (route_unsampled,
src/grower_impl.hpp:430-482).  The device Bernoulli sampler keeps the rows the oracle's
restatement keeps.
"""
import numpy as np
import pytest

from oracle_lib import HIST_F32_FIXED, HIST_INT, cfg as orc_cfg
from test_gpu_depthwise import assert_same_tree, ctx_for, grad_of, tall_data
from test_gpu_growers import LW_F32_CHUNK

pytestmark = pytest.mark.gpu

GROWERS = {"hip_depthwise": "grow_depthwise", "hip_leafwise": "grow_leafwise", "hip_levelwise": "grow_levelwise"}


@pytest.fixture(scope="module")
def tall():
    return tall_data(50000, seed_cfg=4)


def walk(nodes, bins, nb):
    """Leaf of every row by an independent bin-space walk (bin <= split goes left, the
    missing bin goes the default way)."""
    idx = np.zeros(bins.shape[0], np.int64)
    for _ in range(64):
        feat = nodes["feature"][idx]
        inner = feat >= 0
        if not inner.any():
            break
        r = np.nonzero(inner)[0]
        f = feat[r]
        b = bins[r, f].astype(np.int64)
        last = nb[f].astype(np.int64) - 1
        left = np.where(b == last, nodes["default_left"][idx[r]] != 0, b <= nodes["bin"][idx[r]])
        idx[r] = np.where(left, nodes["left"][idx[r]], nodes["right"][idx[r]])
    return idx


def oracle_tree(orc, grower, bins, nb, rows, g, h, gq, hq, sc, quantized, c):
    fn = getattr(orc, GROWERS[grower])
    if quantized:
        return fn(bins[rows], nb, gq=gq[rows], hq=hq[rows], scales=sc, c=c, mode=HIST_INT)
    # the device's fixed-point grids come from every row's gradients, not the subset's
    grids = (orc.fx_shift(g), orc.fx_shift(h))
    if grower == "hip_leafwise":
        return fn(bins[rows], nb, g=g[rows], h=h[rows], scales=grids, c=c, mode=HIST_F32_FIXED, chunk=LW_F32_CHUNK)
    return fn(bins[rows], nb, g=g[rows], h=h[rows], scales=grids, c=c, mode=HIST_F32_FIXED)


def grow_cfg(grower):
    if grower == "hip_leafwise":
        return dict(max_depth=10, max_leaves=31)
    return dict(max_depth=6)


def check_levelwise(nodes, ref):
    # the oblivious grower returns levels, not a node array: compare the dense tree
    levels, lv, lc, gains, rlor = ref
    D = levels.shape[0]
    assert len(nodes) == (2 << D) - 1
    for k in range(D):
        n0 = nodes[(1 << k) - 1]
        assert (n0["feature"], n0["bin"], n0["default_left"]) == tuple(int(x) for x in levels[k])
    leaves = nodes[(1 << D) - 1:]
    assert np.array_equal(leaves["value"].view(np.uint32), lv.view(np.uint32))
    assert np.array_equal(leaves["cover"], lc)


@pytest.mark.parametrize("quantized", [True, False])
@pytest.mark.parametrize("grower", list(GROWERS))
def test_row_subset_tree_and_leaves(hip, orc, tall, grower, quantized):
    bins, nb, X, y = tall
    n = bins.shape[0]
    rng = np.random.default_rng(21)
    rows = np.sort(rng.choice(n, size=int(0.55 * n), replace=False)).astype(np.uint32)
    g, h = grad_of(orc, y, seed=3)
    kw = grow_cfg(grower)
    ctx = ctx_for(hip, bins, nb, grower=grower, quantized=quantized, **kw)
    gq = hq = sc = None
    if quantized:
        gq, hq, sc = ctx.set_gradients(g, h)
    else:
        ctx.set_gradients(g, h)
    ctx.set_row_subset(rows)
    nodes, lor = ctx.grow_tree()
    ctx.close()
    ref = oracle_tree(orc, grower, bins, nb, rows, g, h, gq, hq, sc, quantized, orc_cfg(**kw))
    if grower == "hip_levelwise":
        check_levelwise(nodes, ref)
    else:
        assert_same_tree(nodes, ref[0])
        assert np.array_equal(lor[rows], ref[1])
    assert nodes["cover"][0] == rows.size
    assert np.array_equal(lor, walk(nodes, bins, nb))


def test_row_subset_of_every_row_is_the_plain_root(hip, orc, tall):
    bins, nb, X, y = tall
    g, h = grad_of(orc, y, seed=5)
    out = []
    for subset in (None, np.arange(bins.shape[0], dtype=np.uint32)):
        ctx = ctx_for(hip, bins, nb, grower="hip_depthwise", quantized=True, max_depth=5)
        ctx.set_gradients(g, h)
        ctx.set_row_subset(subset)
        out.append(ctx.grow_tree())
        ctx.close()
    assert np.array_equal(out[0][0].view(np.uint8), out[1][0].view(np.uint8))
    assert np.array_equal(out[0][1], out[1][1])


def test_row_subset_rejects_bad_rows(hip, tall):
    bins, nb, X, y = tall
    ctx = ctx_for(hip, bins, nb, grower="hip_depthwise", quantized=True, max_depth=3)
    for bad in ([5, 3], [1, 1], [bins.shape[0]]):
        with pytest.raises(hip.ConfigError):
            ctx.set_row_subset(np.array(bad, np.uint32))
    with pytest.raises(hip.ConfigError):
        ctx.set_bernoulli(0.0)
    ctx.close()


@pytest.mark.parametrize("quantized", [True, False])
@pytest.mark.parametrize("grower", list(GROWERS))
def test_bernoulli_keeps_the_oracles_rows(hip, orc, tall, grower, quantized):
    bins, nb, X, y = tall
    n = bins.shape[0]
    g, h = grad_of(orc, y, seed=9)
    kw = grow_cfg(grower)
    ctx = ctx_for(hip, bins, nb, grower=grower, quantized=quantized, **kw)
    gq = hq = sc = None
    if quantized:
        gq, hq, sc = ctx.set_gradients(g, h)
    else:
        ctx.set_gradients(g, h)
    ctx.set_bernoulli(0.35, seed=77)
    nodes, lor = ctx.grow_tree()  # a fresh fit: iteration 0 keys the draw
    ctx.close()
    rows = orc.bernoulli_rows(n, 0.35, 77, 0)
    ref = oracle_tree(orc, grower, bins, nb, rows, g, h, gq, hq, sc, quantized, orc_cfg(**kw))
    if grower == "hip_levelwise":
        check_levelwise(nodes, ref)
    else:
        assert_same_tree(nodes, ref[0])
        assert np.array_equal(lor[rows], ref[1])
    assert np.array_equal(lor, walk(nodes, bins, nb))


@pytest.mark.parametrize("sampler", ["rows", "bernoulli"])
def test_sampled_training_graph_equals_eager(hip, orc, tall, sampler):
    bins, nb, X, y = tall
    res = []
    for use_graph in (True, False):
        ctx = ctx_for(hip, bins, nb, grower="hip_depthwise", quantized=True, max_depth=6, use_graph=use_graph,
                      learning_rate=0.1)
        ctx.upload_labels(y)
        ctx.set_scores(np.full(y.size, orc.init_score(1, y), np.float32))
        if sampler == "rows":
            ctx.set_row_subset(np.arange(0, y.size, 3, dtype=np.uint32))
        else:
            ctx.set_bernoulli(0.5, seed=4)
        ctx.train(4)
        res.append((ctx.trees(), ctx.get_scores()))
        ctx.close()
    (ta, sa), (tb, sb) = res
    assert len(ta) == len(tb) == 4
    for a, b in zip(ta, tb):
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8))
    assert np.array_equal(sa.view(np.uint32), sb.view(np.uint32))
    if sampler == "bernoulli":
        # a fresh draw every iteration: the root covers differ
        covers = [t["cover"][0] for t in ta]
        assert len(set(covers)) > 1


def goss_host(g, h, top_rate, other_rate, seed):
    """GossSampler::sample (src/sampler.cpp:73-125) on the host: the top_k rows by |grad|
    kept, other_k drawn uniformly from the rest and their grad/hess amplified by
    (1 - top_rate) / other_rate, rows emitted ascending.  (Ties in |grad| at the cut
    and the uniform draw use numpy's order and generator, not nth_element / std::sample
    with std::mt19937: the seam under test takes whatever rows the sampler chose.)"""
    n = g.size
    top_k, other_k = int(round(top_rate * n)), int(round(other_rate * n))
    order = np.argsort(-np.abs(g), kind="stable")
    rest = order[top_k:]
    picked = np.random.default_rng(seed).choice(rest, size=other_k, replace=False)
    amp = np.float32((np.float32(1.0) - np.float32(top_rate)) / np.float32(other_rate))
    g2, h2 = g.copy(), h.copy()
    g2[picked] *= amp
    h2[picked] *= amp
    rows = np.sort(np.concatenate([order[:top_k], picked])).astype(np.uint32)
    return rows, g2, h2


@pytest.mark.parametrize("quantized", [True, False])
def test_goss_through_the_engine_seam(hip, orc, tall, quantized):
    # GOSS reads gradients, so the reference keeps it off the resident path
    # (include/bonsai/booster.hpp:632-635): the host samples and amplifies, the engine
    # grows on the kept rows with the amplified gradients.
    bins, nb, X, y = tall
    g, h = grad_of(orc, y, seed=13)
    rows, g2, h2 = goss_host(g, h, 0.2, 0.1, seed=6)
    ctx = ctx_for(hip, bins, nb, grower="hip_depthwise", quantized=quantized, max_depth=6)
    gq = hq = sc = None
    if quantized:
        gq, hq, sc = ctx.set_gradients(g2, h2)
    else:
        ctx.set_gradients(g2, h2)
    ctx.set_row_subset(rows)
    nodes, lor = ctx.grow_tree()
    ctx.close()
    ref = oracle_tree(orc, "hip_depthwise", bins, nb, rows, g2, h2, gq, hq, sc, quantized, orc_cfg(max_depth=6))
    assert_same_tree(nodes, ref[0])
    assert np.array_equal(lor[rows], ref[1])
    assert np.array_equal(lor, walk(nodes, bins, nb))


@pytest.mark.parametrize("grower", list(GROWERS))
def test_bernoulli_keeping_no_rows_gives_a_single_leaf(hip, orc, tall, grower):
    # BernoulliSampler can select no row at all; the tree is then one leaf over no rows
    # (zero sums: value 0) and every row lands on it.
    bins, nb, X, y = tall
    n = bins.shape[0]
    assert orc.bernoulli_rows(n, 1e-9, 5, 0).size == 0
    g, h = grad_of(orc, y, seed=2)
    ctx = ctx_for(hip, bins, nb, grower=grower, quantized=True, **grow_cfg(grower))
    ctx.set_gradients(g, h)
    ctx.set_bernoulli(1e-9, seed=5)
    nodes, lor = ctx.grow_tree()
    ctx.close()
    assert nodes[0]["feature"] < 0
    assert nodes[0]["value"] == 0.0 and nodes[0]["cover"] == 0.0
    assert np.all(lor == 0)


def oracle_tree_masked(orc, mask, *a):
    orc.set_feature_mask(mask)
    try:
        return oracle_tree(orc, *a)
    finally:
        orc.set_feature_mask(None)


def check_tree(grower, nodes, lor, ref, rows=None):
    if grower == "hip_levelwise":
        check_levelwise(nodes, ref)
    else:
        assert_same_tree(nodes, ref[0])
        assert np.array_equal(lor if rows is None else lor[rows], ref[1])


@pytest.mark.parametrize("quantized", [True, False])
@pytest.mark.parametrize("grower", list(GROWERS))
def test_feature_mask_matches_oracle(hip, orc, tall, grower, quantized):
    """A handed-over per-tree selection (the `selected` list of sample_features): the
    finders skip the rest, feature 0 included, and the tree equals the oracle's."""
    bins, nb, X, y = tall
    n, F = bins.shape
    mask = np.zeros(F, np.uint8)
    mask[[2, 5, 6, 11, 17, 20, 27]] = 1
    g, h = grad_of(orc, y, seed=13)
    kw = grow_cfg(grower)
    ctx = ctx_for(hip, bins, nb, grower=grower, quantized=quantized, **kw)
    gq = hq = sc = None
    if quantized:
        gq, hq, sc = ctx.set_gradients(g, h)
    else:
        ctx.set_gradients(g, h)
    ctx.set_feature_mask(mask)
    nodes, lor = ctx.grow_tree()
    ctx.close()
    rows = np.arange(n, dtype=np.uint32)
    ref = oracle_tree_masked(orc, mask, grower, bins, nb, rows, g, h, gq, hq, sc, quantized, orc_cfg(**kw))
    check_tree(grower, nodes, lor, ref)
    used = nodes["feature"][nodes["feature"] >= 0]
    assert used.size > 0 and np.all(mask[used] == 1)


@pytest.mark.parametrize("grower", list(GROWERS))
def test_feature_fraction_first_tree_matches_oracle(hip, orc, tall, grower):
    bins, nb, X, y = tall
    n, F = bins.shape
    g, h = grad_of(orc, y, seed=17)
    kw = grow_cfg(grower)
    ctx = ctx_for(hip, bins, nb, grower=grower, quantized=True, **kw)
    gq, hq, sc = ctx.set_gradients(g, h)
    ctx.set_feature_fraction(0.3, seed=41)
    nodes, lor = ctx.grow_tree()  # a fresh fit: iteration 0 keys the draw
    ctx.close()
    mask = orc.feature_draw(F, 0.3, 41, 0)
    assert int(mask.sum()) == 9
    rows = np.arange(n, dtype=np.uint32)
    ref = oracle_tree_masked(orc, mask, grower, bins, nb, rows, g, h, gq, hq, sc, True, orc_cfg(**kw))
    check_tree(grower, nodes, lor, ref)


def test_feature_fraction_and_rows_together(hip, orc, tall):
    """Both samplers at once: the Bernoulli rows and the drawn features of iteration 0."""
    bins, nb, X, y = tall
    n, F = bins.shape
    g, h = grad_of(orc, y, seed=19)
    kw = grow_cfg("hip_leafwise")
    ctx = ctx_for(hip, bins, nb, grower="hip_leafwise", quantized=False, **kw)
    ctx.set_gradients(g, h)
    ctx.set_bernoulli(0.6, seed=8)
    ctx.set_feature_fraction(0.5, seed=9)
    nodes, lor = ctx.grow_tree()
    ctx.close()
    rows = orc.bernoulli_rows(n, 0.6, 8, 0)
    mask = orc.feature_draw(F, 0.5, 9, 0)
    ref = oracle_tree_masked(orc, mask, "hip_leafwise", bins, nb, rows, g, h, None, None, None, False, orc_cfg(**kw))
    check_tree("hip_leafwise", nodes, lor, ref, rows)
    assert np.array_equal(lor, walk(nodes, bins, nb))


@pytest.mark.parametrize("grower", ["hip_depthwise", "hip_levelwise"])
def test_feature_fraction_training_draws_per_tree(hip, orc, tall, grower):
    """A fresh draw every iteration: tree t splits only on the oracle's draw for t, and a
    captured iteration graph trains the same trees as eager launches."""
    bins, nb, X, y = tall
    F = bins.shape[1]
    res = []
    for use_graph in (True, False):
        ctx = ctx_for(hip, bins, nb, grower=grower, quantized=True, max_depth=6, use_graph=use_graph,
                      learning_rate=0.1)
        ctx.upload_labels(y)
        ctx.set_scores(np.full(y.size, orc.init_score(1, y), np.float32))
        ctx.set_feature_fraction(0.25, seed=5)
        ctx.train(3)
        ctx.train(2)
        res.append((ctx.trees(), ctx.get_scores()))
        ctx.close()
    (ta, sa), (tb, sb) = res
    assert len(ta) == len(tb) == 5
    for a, b in zip(ta, tb):
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8))
    assert np.array_equal(sa.view(np.uint32), sb.view(np.uint32))
    sets = []
    for t, tree in enumerate(ta):
        mask = orc.feature_draw(F, 0.25, 5, t)
        used = np.unique(tree["feature"][tree["feature"] >= 0])
        assert used.size > 0 and np.all(mask[used] == 1), t
        sets.append(tuple(used))
    assert len(set(sets)) > 1


def test_feature_sampling_rejects_bad_input(hip, tall):
    bins, nb, X, y = tall
    ctx = ctx_for(hip, bins, nb, grower="hip_depthwise", quantized=True, max_depth=3)
    with pytest.raises(hip.ConfigError):
        ctx.set_feature_mask(np.zeros(bins.shape[1], np.uint8))
    with pytest.raises(hip.ConfigError):
        ctx.set_feature_mask(np.ones(bins.shape[1] + 1, np.uint8))
    with pytest.raises(hip.ConfigError):
        ctx.set_feature_fraction(float("nan"))
    ctx.set_feature_mask(None)
    ctx.set_feature_fraction(1.0)
    ctx.close()


def test_fit_config_subsample_is_the_bernoulli_sampler(hip, orc, tall):
    bins, nb, X, y = tall
    g, h = grad_of(orc, y, seed=43)
    out = []
    for via_config in (True, False):
        kw = dict(subsample=0.45, seed=12) if via_config else dict(seed=12)
        ctx = ctx_for(hip, bins, nb, grower="hip_depthwise", quantized=True, max_depth=5, **kw)
        ctx.set_gradients(g, h)
        if not via_config:
            ctx.set_bernoulli(0.45, seed=12)
        out.append(ctx.grow_tree()[0])
        ctx.close()
    assert out[0].tobytes() == out[1].tobytes()
    assert out[0]["cover"][0] == orc.bernoulli_rows(bins.shape[0], 0.45, 12, 0).size
    with pytest.raises(hip.ConfigError):
        ctx_for(hip, bins, nb, grower="hip_depthwise", quantized=True, max_depth=5, subsample=0.0)
