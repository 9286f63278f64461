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
float gradients within four ulps of max(1, |h|) (Poisson's h = exp(s) scales its
terms); quantized maxima within one float ulp, at least 99.99% of the int8 g codes and
99.9% of the int16 h codes equal (a last-bit difference crosses a floor boundary with
probability about qmax * 2^-24), every other within one quantum.  The oracle's grower then takes the device's own gradients,
and its tree must equal the device's bit for bit; the scores follow the host replay of
s + lr * leaf exactly.
"""
import numpy as np
import pytest

from oracle_lib import CHUNK_ADAPTIVE, HIST_F32_FIXED, HIST_INT, cfg as orc_cfg
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
    """A last-bit difference of exp moves x * qmax / max across a floor boundary with
    probability about qmax * 2^-24: negligible for the int8 g (127 levels), up to ~0.1%
    of rows for the int16 h (32767 levels).  Every code is within one quantum."""
    for k, qmax in ((0, 127.0), (1, 32767.0)):
        md = np.array([sc[k] * qmax], np.float32).view(np.int32)[0]
        mo = np.array([osc[k] * qmax], np.float32).view(np.int32)[0]
        assert abs(int(md) - int(mo)) <= 1, (k, sc, osc)
    for a, b, frac in ((gq, ogq, 0.9999), (hq, ohq, 0.999)):
        d = np.abs(a.astype(np.int32) - b.astype(np.int32))
        assert d.max() <= 1
        assert (d == 0).mean() >= frac, (d == 0).mean()


def oracle_tree(orc, grower, bins, nb, depth, leaves, min_data=20, **grad):
    c = orc_cfg(max_depth=depth, max_leaves=leaves, min_data=min_data)
    if grower == "hip_depthwise":
        mode = HIST_INT if "gq" in grad else HIST_F32_FIXED
        return orc.grow_depthwise(bins, nb, c=c, mode=mode, **grad)[0]
    if grower == "hip_leafwise":
        if "gq" in grad:
            return orc.grow_leafwise(bins, nb, c=c, mode=HIST_INT, **grad)[0]
        return orc.grow_leafwise(bins, nb, c=c, mode=HIST_F32_FIXED, chunk=CHUNK_ADAPTIVE | 256, **grad)[0]
    mode = HIST_INT if "gq" in grad else HIST_F32_FIXED
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


OBJ_ID = {"mse": 0, "logloss": 1, "poisson": 2}


def replay_fit(hip, orc, bins, nb, y, grower, depth, leaves, quantized, iters, objective="logloss", seed=777,
               lr=0.1, min_data=20, check_every=1, weights=None):
    """Train `iters` device iterations one at a time and replay each on the host: the
    device gradients against the oracle's (those through exp within rounding, the
    squared-error ones exactly; weighted rows scaled by w in float), the oracle's tree
    from the device's own gradients against the device tree, the scores bit for bit.
    check_every > 1 checks the gradients and trees of every check_every-th iteration
    only (the scores of all)."""
    lr = np.float32(lr)
    ctx = ctx_for(hip, bins, nb, grower=grower, objective=objective, quantized=quantized, max_depth=depth,
                  max_leaves=leaves, learning_rate=float(lr), seed=seed, min_data_in_leaf=min_data)
    ctx.upload_labels(y)
    if weights is not None:
        ctx.upload_weights(weights)
    init = np.float32(orc.init_score(OBJ_ID[objective], y))
    s = np.full(y.size, init, np.float32)
    ctx.set_scores(s)
    grad_fn = {"logloss": orc.logloss_grad, "mse": orc.mse_grad, "poisson": orc.poisson_grad}[objective]
    exact = objective == "mse"
    for t in range(iters):
        assert np.array_equal(ctx.get_scores().view(np.uint32), s.view(np.uint32)), t
        ctx.train(1)
        tree = ctx.trees()[t]
        if t % check_every == 0:
            g, h = grad_fn(s, y)
            if weights is not None:
                g = (g * weights).astype(np.float32)
                h = (h * weights).astype(np.float32)
            if quantized:
                gq, hq, sc = ctx.get_gradients()
                og, oh, osc = orc.quantize(g, h, seed, t)
                if exact:
                    assert np.array_equal(gq, og) and np.array_equal(hq, oh) and np.array_equal(sc, osc)
                else:
                    assert_codes_agree(gq, hq, sc, og, oh, osc)
                ref = oracle_tree(orc, grower, bins, nb, depth, leaves, min_data, gq=gq, hq=hq, scales=sc)
            else:
                gd, hd = ctx.get_gradients()
                if exact:
                    assert np.array_equal(gd, g) and np.array_equal(hd, h)
                else:
                    # four ulps of the exp-derived terms (|h| scales them), times the weights
                    w = np.ones_like(h) if weights is None else np.maximum(weights, 1.0)
                    tol = 4.0 * 2.0 ** -23 * np.maximum(1.0, np.abs(h)) * w
                    assert np.all(np.abs(gd - g) <= tol) and np.all(np.abs(hd - h) <= tol)
                ref = oracle_tree(orc, grower, bins, nb, depth, leaves, min_data, g=gd, h=hd)
            assert_same_grown(grower, tree, ref)
        s = (s + lr * leaf_values(tree, bins, nb)).astype(np.float32)
    assert np.array_equal(ctx.get_scores().view(np.uint32), s.view(np.uint32))
    trees = ctx.trees()
    ctx.close()
    return trees, s


@pytest.mark.parametrize("quantized", [True, False])
@pytest.mark.parametrize("grower,depth,leaves", [("hip_depthwise", 6, 31), ("hip_leafwise", 10, 63),
                                                 ("hip_levelwise", 5, 31)])
def test_logloss_iterations_match_the_oracle(hip, orc, tall, grower, depth, leaves, quantized):
    """The fused device gradient (and, quantized, the maxima the score update carries to
    the next iteration) against the oracle's logloss_grad + quantize of the replayed
    scores, then the tree grown from the device's own gradients."""
    bins, nb, X, y = tall
    replay_fit(hip, orc, bins, nb, y, grower, depth, leaves, quantized, iters=5)


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


@pytest.mark.parametrize("quantized", [True, False])
@pytest.mark.parametrize("objective", ["logloss", "mse"])
def test_weighted_iterations_match_the_oracle(hip, orc, tall, objective, quantized):
    """Row weights scale the device gradients (gh_from_scores_kernel's Weighted arm,
    src/cuda/detail/kernels.cuh:972-1009): g * w and h * w in float, then the same trees
    as the oracle grown from them."""
    bins, nb, X, y = tall
    w = np.random.default_rng(21).uniform(0.25, 3.0, y.size).astype(np.float32)
    w[::97] = 0.0                                    # zero-weight rows carry no gradient
    target = y if objective == "logloss" else (y * np.float32(3.0) - np.float32(1.0)).astype(np.float32)
    replay_fit(hip, orc, bins, nb, target, "hip_depthwise", 6, 31, quantized, iters=4, objective=objective,
               weights=w)


def poisson_labels(X):
    z = np.where(np.isnan(X[:, 0]), np.float32(0), X[:, 0])
    lam = np.exp(np.float32(0.4) * z + np.float32(0.3) * (X[:, 10] > 0))
    return np.random.default_rng(23).poisson(lam).astype(np.float32)


@pytest.mark.parametrize("quantized", [True, False])
@pytest.mark.parametrize("grower,depth,leaves", [("hip_depthwise", 6, 31), ("hip_leafwise", 10, 63)])
def test_poisson_iterations_match_the_oracle(hip, orc, tall, grower, depth, leaves, quantized):
    """The Poisson device gradient (log link, score clamped to +-30, src/objective.cpp:253-278)."""
    bins, nb, X, y = tall
    replay_fit(hip, orc, bins, nb, poisson_labels(X), grower, depth, leaves, quantized, iters=4, objective="poisson")


def test_weights_refused_and_cleared(hip, orc, tall):
    bins, nb, X, y = tall
    ctx = ctx_for(hip, bins, nb, quantized=True, max_depth=4)
    bad = np.ones(y.size, np.float32)
    bad[5] = -1.0
    with pytest.raises(hip.ConfigError):
        ctx.upload_weights(bad)
    bad[5] = np.inf
    with pytest.raises(hip.ConfigError):
        ctx.upload_weights(bad)
    ctx.upload_labels(y)
    ctx.set_scores(np.zeros(y.size, np.float32))
    ctx.upload_weights(np.full(y.size, 2.0, np.float32))
    ctx.upload_weights(None)                         # cleared: the unweighted fit
    ctx.train(2)
    a = ctx.trees()
    ctx.close()
    ctx = ctx_for(hip, bins, nb, quantized=True, max_depth=4)
    ctx.upload_labels(y)
    ctx.set_scores(np.zeros(y.size, np.float32))
    ctx.train(2)
    b = ctx.trees()
    ctx.close()
    assert all(x.tobytes() == z.tobytes() for x, z in zip(a, b))
