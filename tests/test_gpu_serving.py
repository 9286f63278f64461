"""GPU parity of predict, TreeSHAP and the serving loop.

predict: bit-equal to the oracle's bin-space walk (the reference's contract,
test_cuda_predict.cpp:88-140).  TreeSHAP: within 1e-5 relative of the oracle's
Algorithm 2 in double, additivity sum(phi) = margin within 1e-5 relative
(test_cuda_shap.cpp:55,137-158), golden ice-cream tree against brute-force Shapley.
Serving: every request's result identical alone and inside any batch.
"""
import itertools
import math

import numpy as np
import pytest

from oracle_lib import HIST_F32_OMP, NODE_DTYPE, cfg as orc_cfg
from test_gpu_depthwise import tall_data

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def model():
    from oracle_lib import load_oracle
    orc = load_oracle()
    bins, nb, X, y = tall_data(30000, seed_cfg=4)
    trees, init, _ = orc.train_depthwise(bins[:20000], nb, y[:20000], 1, 40, 0.1, orc_cfg(max_depth=6),
                                         mode=HIST_F32_OMP, n_threads=4)
    return trees, init, bins[20000:], nb


def test_predict_bit_equal(hip, orc, model):
    trees, init, held, nb = model
    p = hip.Predictor(trees, nb, 0.1, init)
    dev = p.predict(held)
    ref = orc.predict(trees, held, nb, 0.1, init)
    assert np.array_equal(dev.view(np.uint32), ref.view(np.uint32))
    # truncated ensembles and a single row
    p5 = hip.Predictor(trees[:5], nb, 0.1, init)
    assert np.array_equal(p5.predict(held[:1]), orc.predict(trees[:5], held[:1], nb, 0.1, init))


def test_shap_matches_algorithm2_and_is_additive(hip, orc, model):
    trees, init, held, nb = model
    rows = held[:257]
    p = hip.Predictor(trees, nb, 0.1, init)
    dev = p.shap(rows)
    ref = orc.pred_contribs(trees, rows, nb, 0.1, init, 6)
    scale = np.abs(ref).sum(axis=1, keepdims=True)
    assert np.max(np.abs(dev - ref) / scale) <= 1e-5
    margin = p.predict(rows).astype(np.float64)
    assert np.max(np.abs(dev.sum(axis=1) - margin) / np.maximum(1.0, np.abs(margin))) <= 1e-5


def ice_cream_tree():
    """The reference's ice-cream tree (tests/unit/test_shap_paths.cpp:421-434):
    covers 100 / 50 50 / 25 25 25 25, leaves 0 10 0 90, both features split at bin 0."""
    t = np.zeros(7, NODE_DTYPE)
    t[0] = (1, 0, 0, 1, 2, 0.0, 100.0, 1.0)
    t[1] = (0, 0, 0, 3, 4, 0.0, 50.0, 1.0)
    t[2] = (0, 0, 0, 5, 6, 0.0, 50.0, 1.0)
    for i, v in zip(range(3, 7), (0.0, 10.0, 0.0, 90.0)):
        t[i] = (-1, 0, 0, 0, 0, v, 25.0, 0.0)
    return t


def brute_shapley(orc, tree, row, nb):
    F = row.size
    last = (np.asarray(nb, np.uint16) - 1).astype(np.uint16)
    import ctypes as C
    from oracle_lib import P
    phi = np.zeros(F)
    tr = np.ascontiguousarray(tree)
    rb = np.ascontiguousarray(row, np.uint8)

    def ev(mask):
        s = np.array([(mask >> f) & 1 for f in range(F)], np.uint8)
        return orc.L.orc_cond_expectation(P(tr), P(rb), P(last), P(s))

    for mask in range(1 << F):
        size = bin(mask).count("1")
        for f in range(F):
            if (mask >> f) & 1:
                continue
            w = math.factorial(size) * math.factorial(F - size - 1) / math.factorial(F)
            phi[f] += w * (ev(mask | (1 << f)) - ev(mask))
    return phi


def test_shap_golden_ice_cream(hip, orc):
    tree = ice_cream_tree()
    nb = np.array([3, 3], np.uint16)   # bins 0, 1 and the missing bin 2
    p = hip.Predictor([tree], nb, 1.0, 0.0)
    rows = np.array(list(itertools.product([0, 1, 2], repeat=2)), np.uint8)
    dev = p.shap(rows)
    for r in range(rows.shape[0]):
        exact = brute_shapley(orc, tree, rows[r], nb)
        assert np.allclose(dev[r, :2], exact, rtol=0, atol=1e-4)
    # bias = expected value of the tree = (0 + 10 + 0 + 90) / 4
    assert np.allclose(dev[:, 2], 25.0)


def test_serve_results_identical_alone_and_in_the_mix(hip, orc, model):
    trees, init, held, nb = model
    p = hip.Predictor(trees, nb, 0.1, init)
    rng = np.random.default_rng(9)
    sizes = np.clip(np.round(rng.lognormal(np.log(64), 1.0, 60)).astype(np.int64), 1, 2000)
    idx = rng.integers(0, held.shape[0], sizes.sum())
    rows = held[idx]
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    arrival = np.cumsum(rng.exponential(1 / 2000.0, sizes.size))
    pred, shap, lat, stats = p.serve(rows, offs, arrival, n_slots=3, do_shap=True, max_batch_rows=4096,
                                     hbm_budget=64 << 20)
    assert stats["requests"] == sizes.size and stats["rows"] == sizes.sum()
    assert np.all(lat > 0) and stats["p99_ms"] >= stats["p50_ms"] > 0
    for r in range(sizes.size):
        a, b = int(offs[r]), int(offs[r + 1])
        alone = p.predict(rows[a:b])
        assert np.array_equal(alone.view(np.uint32), pred[a:b].view(np.uint32))
        sa = p.shap(rows[a:b])
        assert np.array_equal(sa.view(np.uint64), shap[a:b].view(np.uint64))


def test_serve_overlaps_batches_on_the_device(hip, orc, model):
    """Requests that arrive together are admitted into every free slot (several in
    flight at once) and their kernels run on the device at the same time: the device-time
    spans of batches on different streams intersect.  Results stay identical to the
    requests served alone."""
    trees, init, held, nb = model
    p = hip.Predictor(trees, nb, 0.1, init)
    rng = np.random.default_rng(12)
    n_req, size = 16, 6000
    idx = rng.integers(0, held.shape[0], n_req * size)
    rows = held[idx]
    offs = (np.arange(n_req + 1) * size).astype(np.uint64)
    arrival = np.zeros(n_req)                      # a burst: everything arrives at once
    pred, shap, lat, stats = p.serve(rows, offs, arrival, n_slots=4, do_shap=True, max_batch_rows=size,
                                     hbm_budget=1 << 30)
    assert stats["max_inflight"] >= 2
    assert stats["max_device_concurrency"] >= 2
    assert stats["device_overlap_ms"] > 0.0
    for r in (0, 7, n_req - 1):
        a, b = int(offs[r]), int(offs[r + 1])
        assert np.array_equal(p.predict(rows[a:b]).view(np.uint32), pred[a:b].view(np.uint32))
        assert np.array_equal(p.shap(rows[a:b]).view(np.uint64), shap[a:b].view(np.uint64))
    p.close()


def test_shap_split_launch_equals_single_pass(hip, orc, model):
    """A batch of 65,536 or more rows runs one workgroup per 32-row block over every path;
    smaller batches split each block's path groups over several workgroups that add their
    fixed-point sums with 64-bit integer atomics.  Both give the same bits for every row."""
    trees, init, held, nb = model
    p = hip.Predictor(trees, nb, 0.1, init)
    rng = np.random.default_rng(21)
    big = held[rng.integers(0, held.shape[0], 65_536 + 37)]
    whole = p.shap(big)                               # single pass: 2,050 row blocks
    for a, b in ((0, 1), (5, 262), (65_000, 65_573), (1_000, 33_000)):
        part = p.shap(big[a:b])                       # split launches (1 to 1,000 row blocks)
        assert np.array_equal(part.view(np.uint64), whole[a:b].view(np.uint64)), (a, b)
    ref = orc.pred_contribs(trees, big[:64], nb, 0.1, init, 6)
    scale = np.abs(ref).sum(axis=1, keepdims=True)
    assert np.max(np.abs(whole[:64] - ref) / scale) <= 1e-5
    p.close()
