"""The oracle's per-tree feature selection (sample_features, src/grower_impl.hpp:165-180,
with the product's counter-based draw instead of std::sample over std::mt19937) and the
finders' skip of unselected features (src/split.cpp:106-110).

The draw is pinned by what sample_features fixes: exactly max(1, ceil(fraction * F))
distinct features with the product taken in float, every feature for fraction >= 1 or a
single feature, and a uniform spread (over many trees each feature is selected with
probability k / F, checked within 6 standard deviations).  The skip is pinned by an
independent route: a tree grown under a mask equals the tree grown on the selected
columns alone, with the feature ids mapped back.  In FP32 that also pins where node
totals come from: the first selected feature's cells, as SplitInput::totals skips the
unselected placeholders (include/bonsai/split.hpp:36-53); float sums of another
feature's cells differ in the last bits, which the bitwise comparison of gains and leaf
values would show.
"""
import numpy as np
import pytest

from oracle_lib import HIST_F32_CHUNKED, HIST_INT, cfg as orc_cfg


def expected_k(frac, F):
    return max(1, int(np.ceil(np.float32(frac) * np.float32(F))))


@pytest.mark.parametrize("F,frac", [(28, 0.5), (30, 0.1), (10, 0.7), (2000, 0.05), (7, 0.01), (5, 0.0)])
def test_draw_selects_ceil_fraction(orc, F, frac):
    for it in range(5):
        m = orc.feature_draw(F, frac, 9, it)
        assert set(np.unique(m)) <= {0, 1}
        assert int(m.sum()) == expected_k(frac, F)


def test_draw_selects_all_at_or_above_one_or_single_feature(orc):
    for F, frac in ((40, 1.0), (40, 1.5), (1, 0.2)):
        assert np.all(orc.feature_draw(F, frac, 3, 0) == 1)


def test_draw_is_uniform_over_features_and_varies_by_tree(orc):
    F, frac, T = 20, 0.3, 4000
    k = expected_k(frac, F)
    counts = np.zeros(F, np.int64)
    masks = set()
    for it in range(T):
        m = orc.feature_draw(F, frac, 123, it)
        counts += m
        masks.add(m.tobytes())
    p = k / F
    sd = np.sqrt(T * p * (1 - p))
    assert np.all(np.abs(counts - T * p) <= 6 * sd)
    assert len(masks) > T // 2  # C(20, 6) = 38760 possible selections
    # a different seed draws a different sequence
    assert not all(np.array_equal(orc.feature_draw(F, frac, 123, it), orc.feature_draw(F, frac, 124, it))
                   for it in range(8))


def tiny(orc, n=6000, F=12, seed=1):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(n, F)).astype(np.float32)
    y = (X[:, 0] + 0.8 * X[:, 3] - 0.6 * X[:, 7] + 0.3 * rng.normal(size=n) > 0).astype(np.float32)
    cuts = [orc.fit_cuts(X[:, j], 63) for j in range(F)]
    bins = np.stack([orc.transform(cuts[j], X[:, j]) for j in range(F)], axis=1)
    nb = np.array([c.size for c in cuts], np.uint16)
    g, h = orc.mse_grad(np.zeros(n, np.float32), y)
    gq, hq, sc = orc.quantize(g, h, 5, 0)
    return bins, nb, gq, hq, sc


def masked(orc, fn, mask, *a, **kw):
    orc.set_feature_mask(mask)
    try:
        return fn(*a, **kw)
    finally:
        orc.set_feature_mask(None)


def float_grads(orc, n, seed=3):
    rng = np.random.default_rng(seed)
    return orc.logloss_grad(rng.normal(0, 1, n).astype(np.float32), (rng.random(n) < 0.5).astype(np.float32))


@pytest.mark.parametrize("mode", [HIST_INT, HIST_F32_CHUNKED])
@pytest.mark.parametrize("grower", ["grow_depthwise", "grow_leafwise"])
def test_masked_tree_equals_tree_on_selected_columns(orc, grower, mode):
    bins, nb, gq, hq, sc = tiny(orc)
    if mode == HIST_INT:
        grads = dict(gq=gq, hq=hq, scales=sc)
    else:
        g, h = float_grads(orc, bins.shape[0])
        g = (g + np.where(bins[:, 0] > 30, np.float32(0.5), np.float32(-0.5))).astype(np.float32)
        grads = dict(g=g, h=h, chunk=1000)
    sel = np.array([1, 4, 5, 7, 9], np.int64)  # feature 0 (the strongest) left out
    mask = np.zeros(bins.shape[1], np.uint8)
    mask[sel] = 1
    c = orc_cfg(max_depth=6, max_leaves=31)
    fn = getattr(orc, grower)
    full, flor = masked(orc, fn, mask, bins, nb, c=c, mode=mode, **grads)
    sub, slor = fn(np.ascontiguousarray(bins[:, sel]), nb[sel], c=c, mode=mode, **grads)
    assert len(full) == len(sub) > 1
    inner = sub["feature"] >= 0
    assert np.array_equal(full["feature"][inner], sel[sub["feature"][inner]])
    assert np.all(full["feature"][~inner] < 0)
    for f in ("bin", "default_left", "left", "right"):
        assert np.array_equal(full[f][inner], sub[f][inner])
    for f in ("value", "cover", "gain"):
        assert np.array_equal(full[f].view(np.uint32), sub[f].view(np.uint32))
    assert np.array_equal(flor, slor)
    # and the mask changed the tree: the unmasked root splits on feature 0
    plain, _ = fn(bins, nb, c=c, mode=mode, **grads)
    assert plain["feature"][0] == 0


@pytest.mark.parametrize("mode", [HIST_INT, HIST_F32_CHUNKED])
def test_masked_levelwise_equals_selected_columns(orc, mode):
    bins, nb, gq, hq, sc = tiny(orc, seed=2)
    if mode == HIST_INT:
        grads = dict(gq=gq, hq=hq, scales=sc)
    else:
        g, h = float_grads(orc, bins.shape[0], seed=4)
        grads = dict(g=g, h=h, chunk=1000)
    sel = np.array([2, 3, 7, 11], np.int64)
    mask = np.zeros(bins.shape[1], np.uint8)
    mask[sel] = 1
    c = orc_cfg(max_depth=5)
    a = masked(orc, orc.grow_levelwise, mask, bins, nb, c=c, mode=mode, **grads)
    b = orc.grow_levelwise(np.ascontiguousarray(bins[:, sel]), nb[sel], c=c, mode=mode, **grads)
    la, lb = a[0].astype(np.int64), b[0].astype(np.int64)
    assert la.shape == lb.shape and la.shape[0] > 0
    assert np.array_equal(la[:, 0], sel[lb[:, 0]])
    assert np.array_equal(la[:, 1:], lb[:, 1:])
    for x, y in zip(a[1:], b[1:]):
        assert np.array_equal(np.asarray(x), np.asarray(y))
