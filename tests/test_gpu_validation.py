"""Validation plane and early stopping on the device (eval_begin / eval_accumulate,
src/cuda/detail/device_context.cu:1666-1760; the fit loop's rule,
src/cli/pipeline.cpp:443-476).

The recorded loss of every round matches the host loss (the reference's objective eval,
src/objective.cpp:60-120, float terms summed in double) of the oracle's prediction with
the trees so far, within 1e-6 relative; early stopping keeps exactly the trees up to the
best round, which are the trees of the same fit without a validation set.
"""
import numpy as np
import pytest

from test_gpu_depthwise import ctx_for, tall_data

pytestmark = pytest.mark.gpu
N_TR, N_VA = 30000, 10000


@pytest.fixture(scope="module")
def data():
    bins, nb, X, y = tall_data(N_TR + N_VA, seed_cfg=6)
    return bins, nb, y


def host_loss(scores, labels, objective):
    s = np.asarray(scores, np.float32)
    y = np.asarray(labels, np.float32)
    if objective == "logloss":
        t = np.maximum(np.float32(0), s) + np.log1p(np.exp(-np.abs(s))) - y * s
        return t.astype(np.float64).sum() / s.size
    if objective == "poisson":   # src/objective.cpp:280-291
        f = np.clip(s, np.float32(-30.0), np.float32(30.0)).astype(np.float64)
        return (np.exp(f) - y.astype(np.float64) * f).sum() / s.size
    return ((s.astype(np.float64) - y.astype(np.float64)) ** 2).sum() / s.size


def fitted(hip, orc, bins, nb, y, objective, quantized, use_graph=True):
    ctx = ctx_for(hip, bins[:N_TR], nb, quantized=quantized, max_depth=5, learning_rate=0.1,
                  objective=objective, use_graph=use_graph)
    ctx.upload_labels(y[:N_TR])
    init = orc.init_score({"mse": 0, "logloss": 1, "poisson": 2}[objective], y[:N_TR])
    ctx.set_scores(np.full(N_TR, init, np.float32))
    return ctx, init


@pytest.mark.parametrize("objective,quantized", [("logloss", True), ("logloss", False), ("mse", True),
                                                 ("poisson", True)])
def test_loss_history_matches_host(hip, orc, data, objective, quantized):
    bins, nb, y = data
    if objective == "poisson":
        y = np.random.default_rng(3).poisson(np.where(y > 0, 2.0, 0.5)).astype(np.float32)
    ctx, init = fitted(hip, orc, bins, nb, y, objective, quantized)
    yv = y[N_TR:]
    ctx.set_validation(bins[N_TR:], yv, init)
    ctx.train(12)
    hist, best, stopped = ctx.eval_history()
    trees = ctx.trees()
    assert len(hist) == 12 and len(trees) == 12 and not stopped
    for k in range(12):
        p = orc.predict(trees[:k + 1], bins[N_TR:], nb, 0.1, init)
        want = host_loss(p, yv, objective)
        assert abs(float(hist[k]) - want) <= 1e-6 * max(1.0, abs(want)), (k, hist[k], want)
    assert best == int(np.argmin(hist))
    ctx.close()


def test_graph_and_eager_record_the_same_losses(hip, orc, data):
    bins, nb, y = data
    out = []
    for use_graph in (True, False):
        ctx, init = fitted(hip, orc, bins, nb, y, "logloss", True, use_graph=use_graph)
        ctx.set_validation(bins[N_TR:], y[N_TR:], init)
        ctx.train(6)
        out.append(ctx.eval_history()[0])
        ctx.close()
    assert np.array_equal(out[0].view(np.uint32), out[1].view(np.uint32))


def test_early_stopping_keeps_the_trees_up_to_the_best_round(hip, orc, data):
    bins, nb, y = data
    rng = np.random.default_rng(0)
    noise = (rng.random(N_VA) < 0.5).astype(np.float32)  # labels unrelated to the rows
    patience = 3
    ctx, init = fitted(hip, orc, bins, nb, y, "logloss", True)
    ctx.set_validation(bins[N_TR:], noise, init, early_stopping_rounds=patience)
    ctx.train(200)
    hist, best, stopped = ctx.eval_history()
    assert stopped
    assert len(hist) == best + patience + 1          # the round that triggered the stop
    assert hist[best] == hist.min() and int(np.argmin(hist)) == best
    assert np.all(hist[best + 1:] >= hist[best])
    trees = ctx.trees()
    assert len(trees) == best + 1
    ctx.train(5)                                      # a stopped fit stays as it is
    assert len(ctx.trees()) == best + 1
    ctx.close()
    ref, _ = fitted(hip, orc, bins, nb, y, "logloss", True)
    ref.train(best + 1)
    for a, b in zip(trees, ref.trees()):
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8))
    ref.close()


def test_set_scores_disarms_and_bad_inputs_raise(hip, orc, data):
    bins, nb, y = data
    ctx, init = fitted(hip, orc, bins, nb, y, "logloss", True)
    with pytest.raises(hip.ConfigError):
        ctx.set_validation(bins[N_TR:], y[N_TR:N_TR + 5], init)
    ctx.set_validation(bins[N_TR:], y[N_TR:], init)
    ctx.set_scores(np.full(N_TR, init, np.float32))
    with pytest.raises(hip.ConfigError):
        ctx.eval_history()
    ctx.close()
