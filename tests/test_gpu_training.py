"""Quantized boosting iterations against a host replay.

One bh_train call grows several trees; iteration t's gradients come from the scores
after t trees, quantized with their max-abs and the counter RNG keyed on t
(Booster::update_one_iter + resident_round, include/bonsai/booster.hpp:500-676).  The
replay recomputes each step independently: the oracle's squared-error gradients of the
host scores (exact float arithmetic on both sides; the logloss gradient goes through
expf, whose device and host results may differ in the last bit), the oracle's
quantizer, the oracle's grower, and the score step s + lr * leaf in float32 over every
row.  Every device tree must equal the replayed one, which also pins how the device
carries each iteration's gradient maxima to the next.
"""
import numpy as np
import pytest

from oracle_lib import HIST_INT, cfg as orc_cfg
from test_gpu_depthwise import assert_same_tree, ctx_for, tall_data

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
