"""Every benchmark workload's settings on the device against the oracle.

- 100,000 x 28 regression: squared error, depthwise depth 6, 256 bins (the whole u8
  range, the missing bin at 255), 100 iterations, lr 0.1: the full fit replayed on the
  host in both precisions, every tree and the final scores bit for bit.
- 16M x 28 binary at depth 8: one quantized tree on all 16,000,000 rows bit-exact
  against the oracle's integer grower (device binning checked against the oracle's
  transform of a sample), and depth-8 training iterations replayed in both precisions.
- leafwise 255 leaves, depth 12, quantized and FP32: training iterations replayed.
- 1,000,000 x 2,000 u8 (64 bins) levelwise depth 10, on 20,000 rows of the same
  generator: training iterations replayed in both precisions.
- serving: 1,000 random depth-8 trees (paths of up to eight features, the short SHAP
  kernel at its bound) and depth-12 ensembles (longer paths, the 16-element kernel):
  predict bit-equal, TreeSHAP within 1e-5 of Algorithm 2 and additive.

Replays follow tests/test_gpu_training.py: the oracle grows each tree from the
device's own gradients (checked against the oracle's gradients first), and scores
follow s + lr * leaf in float32.
"""
import numpy as np
import pytest

from oracle_lib import HIST_INT, cfg as orc_cfg
from test_gpu_depthwise import assert_same_tree, tall_data
from test_gpu_training import replay_fit

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[configs] {msg}", flush=True)


# ------------------------------------------------------------------ 100k x 28, 256 bins
@pytest.fixture(scope="module")
def reg100k():
    from bonsai_amd import synth
    from oracle_lib import load_oracle
    orc = load_oracle()
    X = synth.tall_features(100_000, config_id=0)
    y = synth.regression_target(X)
    cuts = [orc.fit_cuts(X[:, j], 256) for j in range(X.shape[1])]
    bins = np.stack([orc.transform(cuts[j], X[:, j]) for j in range(X.shape[1])], axis=1)
    nb = np.array([c.size for c in cuts], np.uint16)
    return X, y, cuts, bins, nb


def test_cfg0_bins_fill_the_u8_range(hip, reg100k):
    X, y, cuts, bins, nb = reg100k
    assert nb.max() == 256 and int(bins.max()) == 255        # the missing bin is 255
    ctx = hip.HipContext(0)
    ctx.upload_features(X, cuts)
    assert np.array_equal(ctx.download_bins(), bins)
    ctx.close()


@pytest.mark.parametrize("quantized", [True, False])
def test_cfg0_mse_fit_of_100_iterations(hip, orc, reg100k, quantized):
    X, y, cuts, bins, nb = reg100k
    say(f"cfg0 mse 100k x 28, 256 bins, depth 6, 100 iterations, quantized={quantized}")
    trees, s = replay_fit(hip, orc, bins, nb, y, "hip_depthwise", 6, 31, quantized, iters=100, objective="mse",
                          seed=42)
    assert len(trees) == 100
    # the fit learns: training RMSE well below the target's spread
    assert np.sqrt(np.mean((s.astype(np.float64) - y) ** 2)) < 0.5 * np.std(y)


# ------------------------------------------------------------------ 16M x 28 at depth 8
@pytest.fixture(scope="module")
def tall200k():
    return tall_data(200_000, seed_cfg=1)


@pytest.mark.parametrize("quantized", [True, False])
def test_cfg1_depth8_iterations(hip, orc, tall200k, quantized):
    bins, nb, X, y = tall200k
    say(f"cfg1 depth 8 on 200k rows, quantized={quantized}")
    replay_fit(hip, orc, bins, nb, y, "hip_depthwise", 8, 31, quantized, iters=4)


def test_cfg1_16M_rows_quantized_tree(hip, orc):
    from bonsai_amd import binning, synth
    n = 16_000_000
    say("cfg1: generating 16M x 28")
    X = synth.tall_features(n, config_id=1)
    y = synth.binary_label(X, config_id=1)
    cuts = binning.fit_cuts(X, max_bin=255, n_samples=200_000)
    nb = binning.n_bins_of(cuts)
    ctx = hip.HipContext(0)
    ctx.upload_features(X, cuts)
    bins = ctx.download_bins()
    sample = np.random.default_rng(0).integers(0, n, 50_000)
    for j in range(X.shape[1]):
        assert np.array_equal(bins[sample, j], orc.transform(cuts[j], X[sample, j]))
    del X
    ctx.configure(hip.FitConfig(grower="hip_depthwise", quantized=True, max_depth=8, seed=42))
    g, h = orc.logloss_grad(np.full(n, orc.init_score(1, y), np.float32), y)
    gq, hq, sc = ctx.set_gradients(g, h)
    say("cfg1: device tree")
    nodes, lor = ctx.grow_tree()
    ctx.close()
    say("cfg1: oracle tree")
    ref, rlor = orc.grow_depthwise(bins, nb, gq=gq, hq=hq, scales=sc, c=orc_cfg(max_depth=8), mode=HIST_INT)
    assert_same_tree(nodes, ref)
    assert np.array_equal(lor, rlor)
    assert int((ref["feature"] >= 0).sum()) >= 200       # a deep, bushy tree


# ------------------------------------------------------------------ leafwise 255 leaves, depth 12
@pytest.mark.parametrize("quantized", [True, False])
def test_cfg2_leafwise_255_leaves_iterations(hip, orc, tall200k, quantized):
    bins, nb, X, y = tall200k
    say(f"cfg2 leafwise 255 leaves depth 12, quantized={quantized}")
    trees, _ = replay_fit(hip, orc, bins, nb, y, "hip_leafwise", 12, 255, quantized, iters=3)
    assert all(int((t["feature"] < 0).sum()) == 255 for t in trees)


# ------------------------------------------------------------------ wide levelwise depth 10
@pytest.fixture(scope="module")
def wide20k():
    from bonsai_amd.bench_configs import wide_u8
    bins, y = wide_u8(20_000, 2000)
    return bins, np.full(2000, 66, np.uint16), y


@pytest.mark.parametrize("quantized", [True, False])
def test_cfg3_levelwise_depth10_2000_features(hip, orc, wide20k, quantized):
    bins, nb, y = wide20k
    say(f"cfg3 levelwise depth 10, 20k x 2000, quantized={quantized}")
    trees, _ = replay_fit(hip, orc, bins, nb, y, "hip_levelwise", 10, 31, quantized, iters=3, objective="mse",
                          min_data=5)
    assert len(trees[0]) == (2 << 10) - 1                  # all ten levels grown


# ------------------------------------------------------------------ serving ensembles
def path_features(tree):
    """Longest count of distinct features on a root-to-leaf path."""
    best = 0
    stack = [(0, frozenset())]
    while stack:
        i, fs = stack.pop()
        f = int(tree["feature"][i])
        if f < 0:
            best = max(best, len(fs))
            continue
        fs = fs | {f}
        stack.append((int(tree["left"][i]), fs))
        stack.append((int(tree["right"][i]), fs))
    return best


@pytest.fixture(scope="module")
def serve_rows():
    from bonsai_amd import binning, synth
    X = synth.tall_features(20_000, config_id=0, stream=2)
    cuts = binning.fit_cuts(X, max_bin=256, n_samples=20_000)
    return binning.transform(X, cuts), binning.n_bins_of(cuts)


def check_ensemble(hip, orc, trees, nb, rows, lr, init, depth, shap_rows):
    p = hip.Predictor(trees, nb, lr, init)
    dev = p.predict(rows)
    ref = orc.predict(trees, rows, nb, lr, init)
    assert np.array_equal(dev.view(np.uint32), ref.view(np.uint32))
    sub = rows[:shap_rows]
    phi = p.shap(sub)
    want = orc.pred_contribs(trees, sub, nb, lr, init, depth)
    scale = np.maximum(np.abs(want).sum(axis=1, keepdims=True), 1e-30)
    assert np.max(np.abs(phi - want) / scale) <= 1e-5
    margin = p.predict(sub).astype(np.float64)
    assert np.max(np.abs(phi.sum(axis=1) - margin) / np.maximum(1.0, np.abs(margin))) <= 1e-5
    p.close()


def test_cfg4_thousand_depth8_trees(hip, orc, serve_rows):
    from bonsai_amd.bench_configs import random_ensemble
    rows, nb = serve_rows
    trees = random_ensemble(1000, 8, 28, nb)
    assert max(path_features(t) for t in trees) == 8       # paths at the short kernel's bound
    say("cfg4: 1,000 depth-8 trees")
    check_ensemble(hip, orc, trees, nb, rows[:8192], 1.0, 0.0, 8, 48)


def test_shap_depth12_random_trees(hip, orc, serve_rows):
    from bonsai_amd.bench_configs import random_ensemble
    rows, nb = serve_rows
    trees = random_ensemble(12, 12, 28, nb, seed=5)
    assert max(path_features(t) for t in trees) > 8        # the 16-element kernel runs
    check_ensemble(hip, orc, trees, nb, rows[:4096], 1.0, 0.25, 12, 64)


def test_shap_trained_leafwise_depth12_ensemble(hip, orc, tall200k):
    """A trained leafwise ensemble (255 leaves, depth 12) of the training data."""
    bins, nb, X, y = tall200k
    ctx = hip.HipContext(0)
    ctx.upload_bins(bins, nb)
    ctx.configure(hip.FitConfig(grower="hip_leafwise", quantized=True, max_depth=12, max_leaves=255,
                                learning_rate=0.1))
    ctx.upload_labels(y)
    init = float(orc.init_score(1, y))
    ctx.set_scores(np.full(y.size, init, np.float32))
    ctx.train(6)
    trees = ctx.trees()
    ctx.close()
    assert max(path_features(t) for t in trees) > 8
    check_ensemble(hip, orc, trees, nb, bins[:4096], 0.1, init, 12, 64)
