"""GPU parity of the depthwise hot path against the CPU oracle (through the C ABI).

Bar: bit-exact for integer work (quantized histograms, partition order, decisions and
trees of the quantized grower); FP32 histograms within a stated error of the FP64
oracle and bit-identical across runs; split decisions bit-exact given identical
histograms.
"""
import numpy as np
import pytest

from oracle_lib import DEVICE_CHUNK, HIST_F32_FIXED, HIST_F32_SERIAL, HIST_INT, cfg as orc_cfg

pytestmark = pytest.mark.gpu


def tall_data(n, seed_cfg=0, max_bin=255):
    from bonsai_amd import synth
    from oracle_lib import load_oracle
    orc = load_oracle()
    X = synth.tall_features(n, config_id=seed_cfg)
    y = synth.binary_label(X, config_id=seed_cfg)
    cuts = [orc.fit_cuts(X[:, j], max_bin) for j in range(X.shape[1])]
    bins = np.stack([orc.transform(cuts[j], X[:, j]) for j in range(X.shape[1])], axis=1)
    nb = np.array([c.size for c in cuts], np.uint16)
    return bins, nb, X, y


@pytest.fixture(scope="module")
def tall():
    return tall_data(60000)


def ctx_for(hip, bins, nb, **kw):
    ctx = hip.HipContext(0)
    ctx.upload_bins(bins, nb)
    ctx.configure(hip.FitConfig(**kw))
    return ctx


def grad_of(orc, y, seed=1):
    rng = np.random.default_rng(seed)
    s = rng.normal(0, 1, y.size).astype(np.float32)
    return orc.logloss_grad(s, y)


def test_quantizer_bit_exact(hip, orc, tall):
    bins, nb, X, y = tall
    g, h = grad_of(orc, y)
    ctx = ctx_for(hip, bins, nb, quantized=True, max_depth=6, seed=42)
    gq, hq, sc = ctx.set_gradients(g, h, it=3)
    ogq, ohq, osc = orc.quantize(g, h, 42, 3)
    assert np.array_equal(gq, ogq)
    assert np.array_equal(hq, ohq)
    assert np.array_equal(sc, osc)
    # stochastic rounding stays within one quantum of the value
    assert np.all(np.abs(gq * sc[0] - g) <= sc[0] * (1 + 1e-6))


@pytest.mark.parametrize("frac", [1.0, 0.37, 0.0001])
def test_histogram_int_exact(hip, orc, tall, frac):
    bins, nb, X, y = tall
    g, h = grad_of(orc, y)
    ctx = ctx_for(hip, bins, nb, quantized=True, max_depth=6)
    gq, hq, sc = ctx.set_gradients(g, h)
    rng = np.random.default_rng(5)
    rows = np.flatnonzero(rng.random(y.size) < frac).astype(np.uint32)
    if rows.size == 0:
        rows = np.array([7], np.uint32)
    dev = ctx.histogram(rows)
    ref = orc.hist_int(bins, rows, gq, hq, int(nb.max()))
    assert np.array_equal(dev, ref)
    # every feature sums to the node totals
    tot = ref[0].sum(axis=0)
    assert np.all(dev.sum(axis=1) == tot)


def test_histogram_f32_fixed_order(hip, orc, tall):
    bins, nb, X, y = tall
    g, h = grad_of(orc, y)
    ctx = ctx_for(hip, bins, nb, quantized=False, max_depth=6)
    ctx.set_gradients(g, h)
    rows = np.arange(y.size, dtype=np.uint32)
    d1 = ctx.histogram(rows)
    d2 = ctx.histogram(rows)
    assert np.array_equal(d1.view(np.uint32), d2.view(np.uint32)), "FP32 histogram not deterministic"
    ref64 = orc.hist_f64(bins, rows, g, h, int(nb.max()))
    # error bound of a float sum of n terms: n * eps * sum|x| per cell
    absg = orc.hist_f64(bins, rows, np.abs(g), np.abs(h), int(nb.max()))
    err = np.abs(d1.astype(np.float64) - ref64)
    assert np.all(err <= rows.size * 2 ** -24 * absg + 1e-30)
    chunked = orc.hist_f32(bins, rows, g, h, int(nb.max()), HIST_F32_FIXED, DEVICE_CHUNK)
    # The device sums each DEVICE_CHUNK-row unit exactly on a fixed-point grid and rounds
    # each cell once to float, units folded in order: the FIXED oracle restates that.
    assert np.array_equal(d1.view(np.uint32), chunked.view(np.uint32))


def test_find_split_parity_f32(hip, orc, tall):
    bins, nb, X, y = tall
    g, h = grad_of(orc, y)
    ctx = ctx_for(hip, bins, nb, quantized=False, max_depth=6)
    ctx.set_gradients(g, h)
    c = orc_cfg(max_depth=6)
    rng = np.random.default_rng(11)
    for trial in range(12):
        frac = [1.0, 0.5, 0.05, 0.002][trial % 4]
        rows = np.flatnonzero(rng.random(y.size) < frac).astype(np.uint32)
        hist = ctx.histogram(rows)
        tg, th = orc.totals_f32(hist[0], int(nb[0]))
        ref = orc.find_node_f32(hist, nb, tg, th, rows.size, c)
        dev = ctx.find_split(hist, rows.size)
        assert (dev["valid"], dev["feature"], dev["bin"], dev["default_left"]) == \
            (ref.valid, ref.feature, ref.bin, ref.default_left)
        assert dev["gain"] == ref.gain


def test_find_split_parity_int(hip, orc, tall):
    bins, nb, X, y = tall
    g, h = grad_of(orc, y)
    ctx = ctx_for(hip, bins, nb, quantized=True, max_depth=6)
    gq, hq, sc = ctx.set_gradients(g, h)
    c = orc_cfg(max_depth=6)
    rng = np.random.default_rng(12)
    for trial in range(12):
        frac = [1.0, 0.5, 0.05, 0.002][trial % 4]
        rows = np.flatnonzero(rng.random(y.size) < frac).astype(np.uint32)
        hist = ctx.histogram(rows)
        tg, th = hist[0, :, 0].sum(), hist[0, :, 1].sum()
        ref = orc.find_node_int(hist, nb, tg, th, sc[0], sc[1], rows.size, c)
        dev = ctx.find_split(hist, rows.size, scales=sc)
        assert (dev["valid"], dev["feature"], dev["bin"], dev["default_left"]) == \
            (ref.valid, ref.feature, ref.bin, ref.default_left)
        assert dev["gain"] == ref.gain


@pytest.mark.parametrize("feature,bin_,dl", [(0, 100, 1), (5, 3, 0), (20, 40, 1), (25, 7, 0), (27, 0, 1),
                                              (3, 254, 0)])
def test_partition_parity(hip, orc, tall, feature, bin_, dl):
    bins, nb, X, y = tall
    ctx = ctx_for(hip, bins, nb, quantized=False, max_depth=6)
    ctx.set_gradients(np.zeros(y.size, np.float32), np.ones(y.size, np.float32))
    rng = np.random.default_rng(feature)
    rows = np.flatnonzero(rng.random(y.size) < 0.6).astype(np.uint32)
    dev, nl = ctx.partition(rows, feature, bin_, bool(dl))
    ref, rnl = orc.partition(bins, rows, feature, bin_, dl, int(nb[feature]) - 1)
    assert nl == rnl
    assert np.array_equal(dev, ref)


TREE_FIELDS = ("feature", "bin", "default_left", "left", "right", "value", "cover", "gain")


def assert_same_tree(dev, ref):
    assert len(dev) == len(ref)
    for f in TREE_FIELDS:
        a, b = dev[f], ref[f]
        leaf = ref["feature"] < 0
        if f in ("bin", "default_left", "left", "right"):
            a, b = a[~leaf], b[~leaf]
        assert np.array_equal(a, b), f"field {f} differs"


@pytest.mark.parametrize("pair", [1, 0])
@pytest.mark.parametrize("depth", [1, 2, 4, 7])
def test_grow_depthwise_int_bit_exact(hip, orc, tall, depth, pair):
    """pair=1 moves rows every second level (four-way partitions, routed histogram
    units in between); pair=0 moves them after every split."""
    bins, nb, X, y = tall
    g, h = grad_of(orc, y, seed=depth)
    ctx = ctx_for(hip, bins, nb, quantized=True, max_depth=depth, min_data_in_leaf=20)
    ctx.set_tuning(pair_levels=pair)
    gq, hq, sc = ctx.set_gradients(g, h)
    nodes, lor = ctx.grow_tree()
    ref, rlor = orc.grow_depthwise(bins, nb, gq=gq, hq=hq, scales=sc, c=orc_cfg(max_depth=depth), mode=HIST_INT)
    assert_same_tree(nodes, ref)
    assert np.array_equal(lor, rlor)


@pytest.mark.parametrize("width", [10, 60])
def test_grow_depthwise_int_paired_row_widths(hip, orc, tall, width):
    """Rows of 12 and 60 bytes through the four-way partitions, whose tiles reserve
    their class runs with atomic cursors (4,096 workgroups racing), give the oracle's
    tree and leaf of every row."""
    bins, nb, X, y = tall
    reps = -(-width // bins.shape[1])
    wb = np.ascontiguousarray(np.hstack([bins] * reps)[:, :width])
    wnb = np.concatenate([nb] * reps)[:width]
    g, h = grad_of(orc, y, seed=width)
    ctx = ctx_for(hip, wb, wnb, quantized=True, max_depth=6, min_data_in_leaf=20)
    ctx.set_tuning(pair_levels=1, part_grid=4096)
    gq, hq, sc = ctx.set_gradients(g, h)
    nodes, lor = ctx.grow_tree()
    ref, rlor = orc.grow_depthwise(wb, wnb, gq=gq, hq=hq, scales=sc, c=orc_cfg(max_depth=6), mode=HIST_INT)
    assert_same_tree(nodes, ref)
    assert np.array_equal(lor, rlor)


@pytest.mark.parametrize("depth", [3, 6])
def test_grow_depthwise_f32_matches_device_order_oracle(hip, orc, tall, depth):
    bins, nb, X, y = tall
    g, h = grad_of(orc, y, seed=10 + depth)
    ctx = ctx_for(hip, bins, nb, quantized=False, max_depth=depth)
    ctx.set_gradients(g, h)
    nodes, lor = ctx.grow_tree()
    ref, rlor = orc.grow_depthwise(bins, nb, g=g, h=h, c=orc_cfg(max_depth=depth), mode=HIST_F32_FIXED)
    assert_same_tree(nodes, ref)
    assert np.array_equal(lor, rlor)
    # against the serial reference fill: same routing within the reference's 1e-4 band
    ser, slor = orc.grow_depthwise(bins, nb, g=g, h=h, c=orc_cfg(max_depth=depth), mode=HIST_F32_SERIAL)
    vd = nodes["value"][lor]
    vs = ser["value"][slor]
    assert np.max(np.abs(vd - vs)) <= 1e-4


@pytest.mark.parametrize("quantized", [False, True])
def test_train_graph_equals_eager_and_is_deterministic(hip, orc, tall, quantized):
    bins, nb, X, y = tall
    res = []
    for use_graph in (True, False, True):
        ctx = ctx_for(hip, bins, nb, quantized=quantized, max_depth=5, use_graph=use_graph,
                      learning_rate=0.1, objective="logloss")
        ctx.upload_labels(y)
        init = orc.init_score(1, y)
        ctx.set_scores(np.full(y.size, init, np.float32))
        ctx.train(4)
        ctx.train(3)
        res.append((ctx.trees(), ctx.get_scores()))
        ctx.close()
    for t, s in res[1:]:
        assert len(t) == len(res[0][0]) == 7
        for a, b in zip(t, res[0][0]):
            assert np.array_equal(a.view(np.uint8), b.view(np.uint8))
        assert np.array_equal(s.view(np.uint32), res[0][1].view(np.uint32))


def test_train_f32_matches_oracle_fit(hip, orc, tall):
    """Whole fit: device gradients use the device expf, so trees are compared as
    predictions within the reference's CPU-vs-GPU band (test_cuda_grower.cpp:94)."""
    bins, nb, X, y = tall
    ctx = ctx_for(hip, bins, nb, quantized=False, max_depth=4, learning_rate=0.1)
    ctx.upload_labels(y)
    init = orc.init_score(1, y)
    ctx.set_scores(np.full(y.size, init, np.float32))
    ctx.train(5)
    dev_scores = ctx.get_scores()
    trees, oinit, oscores = orc.train_depthwise(bins, nb, y, 1, 5, 0.1, orc_cfg(max_depth=4), mode=HIST_F32_SERIAL,
                                                n_threads=4)
    assert oinit == init
    assert np.max(np.abs(dev_scores - oscores)) <= 1e-4 * 5
