"""Boosting iterations against a host replay.

One bh_train call grows several trees; iteration t's gradients come from the scores
after t trees, quantized with their max-abs and the counter RNG keyed on t
(Booster::update_one_iter + resident_round, include/bonsai/booster.hpp:500-676).  The
replay recomputes each step independently: the oracle's squared-error gradients of the
host scores (exact float arithmetic on both sides), the oracle's quantizer, the oracle's grower, and the score step s + lr * leaf in float32 over every
row.  Every device tree must equal the replayed one, which also pins how the device
carries each iteration's gradient maxima to the next.

The logistic objective goes through expf on the device and std::exp in the oracle,
which may differ in the last bit, so its replay reads each iteration's device gradients
back (bh_get_gradients) and checks them against the oracle's within that rounding:
float gradients within 2e-7 absolute (a few ulps of values in [-1, 1]); quantized
maxima within one float ulp, at least 99.99% of the int8 / int16 codes equal and every
other within one quantum.  The oracle's grower then takes the device's own gradients,
and its tree must equal the device's bit for bit; the scores follow the host replay of
s + lr * leaf exactly.
"""
import numpy as np
import pytest

from oracle_lib import CHUNK_ADAPTIVE, HIST_F32_CHUNKED, HIST_INT, cfg as orc_cfg
from test_gpu_depthwise import assert_same_tree, ctx_for, tall_data
from test_gpu_growers import levelwise_view

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tall():
    return tall_data(60000, seed_cfg=5)


def leaf_values(nodes, bins, nb):
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
    return nodes["value"][idx]


@pytest.mark.parametrize("calls", [[5], [2, 3], [1, 1, 1, 1, 1]])
def test_quantized_iterations_match_the_host_replay(hip, orc, tall, calls):
    bins, nb, X, y = tall
    lr, seed, depth = np.float32(0.1), 1234, 5
    target = (y * np.float32(2.0) + X[:, 0].astype(np.float32) * np.float32(0.25)).astype(np.float32)
    target = np.where(np.isnan(target), np.float32(0.0), target).astype(np.float32)
    ctx = ctx_for(hip, bins, nb, grower="hip_depthwise", objective="mse", quantized=True, max_depth=depth,
                  learning_rate=float(lr), seed=seed)
    ctx.upload_labels(target)
    init = np.float32(orc.init_score(0, target))
    ctx.set_scores(np.full(target.size, init, np.float32))
    for k in calls:
        ctx.train(k)
    trees = ctx.trees()
    dev_scores = ctx.get_scores()
    ctx.close()
    assert len(trees) == sum(calls)
    s = np.full(target.size, init, np.float32)
    for t, tree in enumerate(trees):
        g, h = orc.mse_grad(s, target)
        gq, hq, sc = orc.quantize(g, h, seed, t)
        ref, _ = orc.grow_depthwise(bins, nb, gq=gq, hq=hq, scales=sc, c=orc_cfg(max_depth=depth), mode=HIST_INT)
        assert_same_tree(tree, ref)
        s = (s + lr * leaf_values(tree, bins, nb)).astype(np.float32)
    assert np.array_equal(dev_scores.view(np.uint32), s.view(np.uint32))


def assert_codes_agree(gq, hq, sc, ogq, ohq, osc):
    for k, qmax in ((0, 127.0), (1, 32767.0)):
        md = np.array([sc[k] * qmax], np.float32).view(np.int32)[0]
        mo = np.array([osc[k] * qmax], np.float32).view(np.int32)[0]
        assert abs(int(md) - int(mo)) <= 1, (k, sc, osc)
    for a, b in ((gq, ogq), (hq, ohq)):
        d = np.abs(a.astype(np.int32) - b.astype(np.int32))
        assert d.max() <= 1
        assert (d == 0).mean() >= 0.9999


def oracle_tree(orc, grower, bins, nb, depth, leaves, **grad):
    c = orc_cfg(max_depth=depth, max_leaves=leaves)
    if grower == "hip_depthwise":
        mode = HIST_INT if "gq" in grad else HIST_F32_CHUNKED
        return orc.grow_depthwise(bins, nb, c=c, mode=mode, **grad)[0]
    if grower == "hip_leafwise":
        if "gq" in grad:
            return orc.grow_leafwise(bins, nb, c=c, mode=HIST_INT, **grad)[0]
        return orc.grow_leafwise(bins, nb, c=c, mode=HIST_F32_CHUNKED, chunk=CHUNK_ADAPTIVE | 256, **grad)[0]
    mode = HIST_INT if "gq" in grad else HIST_F32_CHUNKED
    return orc.grow_levelwise(bins, nb, c=c, mode=mode, **grad)


def assert_same_grown(grower, tree, ref):
    if grower != "hip_levelwise":
        assert_same_tree(tree, ref)
        return
    levels, lv, lc = ref[0], ref[1], ref[2]
    D = levels.shape[0]
    dl, dv, dc = levelwise_view(tree, D)
    assert np.array_equal(dl, levels.astype(np.int64))
    assert np.array_equal(dv.view(np.uint32), lv.view(np.uint32))
    assert np.array_equal(dc, lc)


@pytest.mark.parametrize("quantized", [True, False])
@pytest.mark.parametrize("grower,depth,leaves", [("hip_depthwise", 6, 31), ("hip_leafwise", 10, 63),
                                                 ("hip_levelwise", 5, 31)])
def test_logloss_iterations_match_the_oracle(hip, orc, tall, grower, depth, leaves, quantized):
    """The fused device gradient (and, quantized, the maxima the score update carries to
    the next iteration) against the oracle's logloss_grad + quantize of the replayed
    scores, then the tree grown from the device's own gradients."""
    bins, nb, X, y = tall
    lr, seed, iters = np.float32(0.1), 777, 5
    ctx = ctx_for(hip, bins, nb, grower=grower, objective="logloss", quantized=quantized, max_depth=depth,
                  max_leaves=leaves, learning_rate=float(lr), seed=seed)
    ctx.upload_labels(y)
    init = np.float32(orc.init_score(1, y))
    s = np.full(y.size, init, np.float32)
    ctx.set_scores(s)
    for t in range(iters):
        assert np.array_equal(ctx.get_scores().view(np.uint32), s.view(np.uint32)), t
        ctx.train(1)
        g, h = orc.logloss_grad(s, y)
        if quantized:
            gq, hq, sc = ctx.get_gradients()
            assert_codes_agree(gq, hq, sc, *orc.quantize(g, h, seed, t))
            ref = oracle_tree(orc, grower, bins, nb, depth, leaves, gq=gq, hq=hq, scales=sc)
        else:
            gd, hd = ctx.get_gradients()
            assert np.max(np.abs(gd - g)) <= 2e-7 and np.max(np.abs(hd - h)) <= 2e-7
            ref = oracle_tree(orc, grower, bins, nb, depth, leaves, g=gd, h=hd)
        tree = ctx.trees()[t]
        assert_same_grown(grower, tree, ref)
        s = (s + lr * leaf_values(tree, bins, nb)).astype(np.float32)
    assert np.array_equal(ctx.get_scores().view(np.uint32), s.view(np.uint32))
    ctx.close()


def test_gradients_read_back_as_handed_over(hip, orc, tall):
    """bh_get_gradients after bh_set_gradients returns what the quantizer made of them."""
    bins, nb, X, y = tall
    rng = np.random.default_rng(9)
    g, h = orc.logloss_grad(rng.normal(0, 1, y.size).astype(np.float32), y)
    ctx = ctx_for(hip, bins, nb, quantized=True, max_depth=4, seed=5)
    gq, hq, sc = ctx.set_gradients(g, h, it=2)
    rq, rh, rs = ctx.get_gradients()
    assert np.array_equal(gq, rq) and np.array_equal(hq, rh) and np.array_equal(sc, rs)
    ctx.close()
    ctx = ctx_for(hip, bins, nb, quantized=False, max_depth=4)
    ctx.set_gradients(g, h)
    rg, rh = ctx.get_gradients()
    assert np.array_equal(rg, g) and np.array_equal(rh, h)
    ctx.close()
