"""The oracle against what the mathematics fixes: brute force on small inputs, closed
forms, conservation laws and special cases, each chosen so that a plausible mistake
(a dropped term, a wrong index or sign, an off-by-one bin) fails at least one check.
"""
import ctypes as C
import itertools
import math
from fractions import Fraction

import numpy as np
import pytest

from oracle_lib import (CHUNK_ADAPTIVE, HIST_F32_CHUNKED, HIST_F32_FIXED, HIST_F32_SERIAL, HIST_INT, NODE_DTYPE, P,
                        TreeCfg, cfg as orc_cfg)


def small_plane(n=3000, f=5, nb=12, seed=0, miss=0.05):
    rng = np.random.default_rng(seed)
    bins = rng.integers(0, nb - 1, (n, f)).astype(np.uint8)
    bins[rng.random((n, f)) < miss] = nb - 1      # the missing bin is the last
    return bins, np.full(f, nb, np.uint16)


# ----------------------------------------------------------------------------- histograms
def test_hist_brute_force_and_conservation(orc):
    bins, nb = small_plane()
    rng = np.random.default_rng(1)
    g = rng.normal(0, 1, bins.shape[0]).astype(np.float32)
    h = rng.uniform(0.1, 1, bins.shape[0]).astype(np.float32)
    rows = np.sort(rng.choice(bins.shape[0], 1700, replace=False)).astype(np.uint32)
    H = orc.hist_f64(bins, rows, g, h, 12)
    brute = np.zeros_like(H)
    for r in rows:
        for j in range(bins.shape[1]):
            brute[j, bins[r, j], 0] += g[r]
            brute[j, bins[r, j], 1] += h[r]
    assert np.allclose(H, brute, rtol=1e-12, atol=1e-12)
    # every feature's cells sum to the node totals
    tot = np.array([g[rows].astype(np.float64).sum(), h[rows].astype(np.float64).sum()])
    assert np.allclose(H.sum(axis=1), tot[None, :], rtol=1e-12)


def test_hist_int_exact_and_sibling_subtraction(orc):
    bins, nb = small_plane(seed=2)
    rng = np.random.default_rng(3)
    gq = rng.integers(-127, 128, bins.shape[0]).astype(np.int8)
    hq = rng.integers(0, 32768, bins.shape[0]).astype(np.int16)
    rows = np.arange(bins.shape[0], dtype=np.uint32)
    parent = orc.hist_int(bins, rows, gq, hq, 12)
    left_rows, nl = orc.partition(bins, rows, 2, 4, True, 11)
    L = orc.hist_int(bins, left_rows[:nl], gq, hq, 12)
    R = orc.hist_int(bins, left_rows[nl:], gq, hq, 12)
    assert np.array_equal(parent, L + R)          # exact in integers
    assert np.array_equal(parent.sum(axis=1)[0], [int(gq.astype(np.int64).sum()), int(hq.astype(np.int64).sum())])


def test_hist_f32_single_chunk_is_serial_order_and_error_bound(orc):
    """test_populate.cpp:160-182: one block reproduces the serial order bit for bit.
    Across chunks the float result stays within n * eps of the FP64 oracle."""
    bins, nb = small_plane(n=5000, seed=4)
    rng = np.random.default_rng(5)
    g = rng.normal(0, 1, bins.shape[0]).astype(np.float32)
    h = rng.uniform(0.1, 1, bins.shape[0]).astype(np.float32)
    rows = np.arange(bins.shape[0], dtype=np.uint32)
    serial = orc.hist_f32(bins, rows, g, h, 12, HIST_F32_SERIAL)
    one = orc.hist_f32(bins, rows, g, h, 12, HIST_F32_CHUNKED, chunk=1 << 20)
    assert np.array_equal(serial.view(np.uint32), one.view(np.uint32))
    many = orc.hist_f32(bins, rows, g, h, 12, HIST_F32_CHUNKED, chunk=512)
    ref = orc.hist_f64(bins, rows, g, h, 12)
    absum = orc.hist_f64(bins, rows, np.abs(g), h, 12)
    bound = 5000 * np.finfo(np.float32).eps * absum
    assert np.all(np.abs(many - ref) <= bound + 1e-30)


# ------------------------------------------------------- FP32 fixed-point unit sums
def _f32_of_scaled(q: int, s: int) -> np.float32:
    """q * 2^-s rounded once to float32 (nearest, ties to even), by integer arithmetic."""
    if q == 0:
        return np.float32(0.0)
    a = abs(q)
    shift = max(0, a.bit_length() - 24)
    m, rem = a >> shift, a & ((1 << shift) - 1)
    half = 1 << (shift - 1) if shift else 0
    if shift and (rem > half or (rem == half and (m & 1))):
        m += 1
    return np.float32(math.copysign(math.ldexp(m, shift - s), q))


def _fixed_cells(bins, rows, v, s, SB):
    """Exact grid sums per (feature, bin): v on the grid 2^-s, rounded half to even."""
    F = bins.shape[1]
    acc = [[0] * SB for _ in range(F)]
    for r in rows:
        q = round(Fraction(float(v[r])) * (2 ** s) if s >= 0 else Fraction(float(v[r])) / (2 ** -s))
        for f in range(F):
            acc[f][int(bins[r, f])] += q
    return acc


def test_fx_shift_is_49_minus_the_max_exponent(orc):
    """s = 49 - e with max|v| < 2^e: 1.0 -> 48, just below 1 -> 49, 0.25 -> 50, 3 -> 47,
    the sign is ignored, an all-zero array gives 0."""
    for vals, want in (([1.0], 48), ([0.999], 49), ([0.25, -0.1], 50), ([-3.0, 2.0], 47), ([0.0, 0.0], 0),
                       ([1e-30], 49 - math.frexp(1e-30)[1]), ([7.5e8], 49 - 30)):
        assert orc.fx_shift(np.array(vals, np.float32)) == want, vals


def test_fixed_unit_is_the_correctly_rounded_exact_sum(orc):
    """One unit: every cell is the exact sum of its rows' grid values rounded once to
    float32 (tiny, huge-ratio and negative values included); several units: the rounded
    unit cells added in unit order in float32.  A dropped rounding step, a wrong grid or
    a fold out of order changes some cell."""
    bins, nb = small_plane(n=2500, f=3, seed=40)
    rng = np.random.default_rng(41)
    n = bins.shape[0]
    g = (rng.normal(0, 1, n) * 10.0 ** rng.uniform(-12, 0, n)).astype(np.float32)
    h = rng.uniform(1e-9, 0.25, n).astype(np.float32)
    rows = np.arange(n, dtype=np.uint32)
    SB = 12
    sg, sh = orc.fx_shift(g), orc.fx_shift(h)
    for chunk in (8192, 1000):
        got = orc.hist_f32(bins, rows, g, h, SB, HIST_F32_FIXED, chunk=chunk)
        want = np.zeros((3, SB, 2), np.float32)
        for k0 in range(0, n, chunk):
            part = rows[k0:k0 + chunk]
            ag, ah = _fixed_cells(bins, part, g, sg, SB), _fixed_cells(bins, part, h, sh, SB)
            for f in range(3):
                for b in range(SB):
                    want[f, b, 0] = np.float32(want[f, b, 0] + _f32_of_scaled(ag[f][b], sg))
                    want[f, b, 1] = np.float32(want[f, b, 1] + _f32_of_scaled(ah[f][b], sh))
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), chunk
    # and it is at least as close to the FP64 sums as the serial float fill
    ref = orc.hist_f64(bins, rows, g, h, SB)
    ser = orc.hist_f32(bins, rows, g, h, SB, HIST_F32_SERIAL)
    one = orc.hist_f32(bins, rows, g, h, SB, HIST_F32_FIXED, chunk=8192)
    assert np.abs(one - ref).sum() <= np.abs(ser - ref).sum()


# ----------------------------------------------------------------------------- quantizer
def test_quantizer_bounds_unbiased_deterministic(orc):
    rng = np.random.default_rng(6)
    g = rng.normal(0, 1, 4000).astype(np.float32)
    h = rng.uniform(0, 0.25, 4000).astype(np.float32)
    gq, hq, sc = orc.quantize(g, h, 42, 0)
    assert sc[0] == pytest.approx(np.abs(g).max() / 127.0) and sc[1] == pytest.approx(np.abs(h).max() / 32767.0)
    assert np.all(np.abs(gq * sc[0] - g) <= sc[0] * (1 + 1e-6))
    assert np.all(np.abs(hq * sc[1] - h) <= sc[1] * (1 + 1e-6))
    assert np.abs(gq).max() <= 127 and hq.min() >= 0
    again = orc.quantize(g, h, 42, 0)
    assert np.array_equal(gq, again[0]) and np.array_equal(hq, again[1])
    # stochastic rounding is unbiased: E[q * scale] = x over independent draws
    acc = np.zeros(g.size)
    T = 400
    for it in range(T):
        acc += orc.quantize(g, h, 7, it)[0] * sc[0]
    err = acc / T - g
    assert np.abs(err.mean()) < 3 * sc[0] / math.sqrt(T * g.size) * 2
    assert np.abs(err).max() < 5 * sc[0] / math.sqrt(T) + 1e-6


# ----------------------------------------------------------------------------- partition
@pytest.mark.parametrize("dl", [True, False])
def test_partition_stable_and_routes_left(orc, dl):
    """routes_left (include/bonsai/dataset.hpp:234-238): bin == last ? dl : bin <= split."""
    bins, nb = small_plane(seed=7)
    rng = np.random.default_rng(8)
    rows = np.sort(rng.choice(bins.shape[0], 2000, replace=False)).astype(np.uint32)
    out, nl = orc.partition(bins, rows, 3, 5, dl, 11)
    go_left = [(dl if bins[r, 3] == 11 else bins[r, 3] <= 5) for r in rows]
    exp = np.concatenate([rows[np.array(go_left)], rows[~np.array(go_left)]])
    assert nl == int(np.sum(go_left))
    assert np.array_equal(out, exp)


# ----------------------------------------------------------------------------- growers
def test_separable_stump(orc):
    """test_grower.cpp: a step target on one feature gives a depth-1 tree at the step with
    leaf values -G/(H + lambda)."""
    n = 400
    bins = np.zeros((n, 3), np.uint8)
    rng = np.random.default_rng(9)
    bins[:, 0] = rng.integers(0, 9, n)
    bins[:, 1] = rng.integers(0, 9, n)
    bins[:, 2] = np.arange(n) % 9
    y = np.where(bins[:, 1] <= 3, -2.0, 3.0).astype(np.float32)
    g, h = orc.mse_grad(np.zeros(n, np.float32), y)
    nb = np.full(3, 10, np.uint16)
    c = orc_cfg(max_depth=1, min_data=1, l2=1.0)
    nodes, lor = orc.grow_depthwise(bins, nb, g=g, h=h, c=c)
    assert len(nodes) == 3
    assert nodes[0]["feature"] == 1 and nodes[0]["bin"] == 3
    left = bins[:, 1] <= 3
    for child, mask in ((nodes[0]["left"], left), (nodes[0]["right"], ~left)):
        G = g[mask].astype(np.float64).sum()
        H = h[mask].astype(np.float64).sum()
        assert nodes[child]["value"] == np.float32(-G / (H + 1.0))
        assert np.all(lor[mask] == child)


def test_l1_soft_threshold_leaf(orc):
    """bounded_leaf_weight: w = -sign(G) max(|G| - l1, 0) / (H + l2)."""
    for G, H, l1, l2 in [(3.0, 2.0, 1.0, 1.0), (-3.0, 2.0, 1.0, 1.0), (0.5, 2.0, 1.0, 1.0), (7.0, 0.0, 0.0, 2.0)]:
        c = TreeCfg(0.0, 0.0, l2, l1, 6, 0, 31, 0)
        t = math.copysign(max(abs(G) - l1, 0.0), G)
        assert orc.leaf_value(G, H, c) == np.float32(-t / (H + l2) if H + l2 > 0 else 0.0)


def test_depthwise_int_equals_f64_when_exact(orc):
    """Integer gradients make every float and fixed-point sum exact: the FP32 serial,
    FP32 chunked and integer growers must build the same tree."""
    bins, nb = small_plane(n=4000, seed=10)
    rng = np.random.default_rng(11)
    gq = rng.integers(-127, 128, bins.shape[0]).astype(np.int8)
    hq = rng.integers(1, 200, bins.shape[0]).astype(np.int16)
    g = gq.astype(np.float32)
    h = hq.astype(np.float32)
    c = orc_cfg(max_depth=4, min_data=5)
    a, la = orc.grow_depthwise(bins, nb, g=g, h=h, c=c, mode=HIST_F32_SERIAL)
    b, lb = orc.grow_depthwise(bins, nb, g=g, h=h, c=c, mode=HIST_F32_CHUNKED, chunk=700)
    x, lx = orc.grow_depthwise(bins, nb, g=g, h=h, c=c, mode=HIST_F32_FIXED, chunk=700)
    q, lq = orc.grow_depthwise(bins, nb, gq=gq, hq=hq, scales=(1.0, 1.0), c=c, mode=HIST_INT)
    for t in (b, x, q):
        assert np.array_equal(a["feature"], t["feature"])
        internal = a["feature"] >= 0
        assert np.array_equal(a["bin"][internal], t["bin"][internal])
        assert np.array_equal(a["value"], t["value"])
    assert np.array_equal(la, lb) and np.array_equal(la, lx) and np.array_equal(la, lq)


def test_levelwise_int_equals_f32_when_exact(orc):
    """The levelwise twin: with integer gradients the FP32 serial and the integer level
    finders see the same exact sums, so the oblivious trees (level decisions, leaf values,
    covers, leaf of every row) must agree."""
    bins, nb = small_plane(n=6000, f=6, seed=30)
    rng = np.random.default_rng(31)
    gq = rng.integers(-127, 128, bins.shape[0]).astype(np.int8)
    hq = rng.integers(1, 200, bins.shape[0]).astype(np.int16)
    g = gq.astype(np.float32)
    h = hq.astype(np.float32)
    for depth in (1, 3, 5):
        c = orc_cfg(max_depth=depth, min_data=5)
        a = orc.grow_levelwise(bins, nb, g=g, h=h, c=c, mode=HIST_F32_SERIAL)
        q = orc.grow_levelwise(bins, nb, gq=gq, hq=hq, scales=(1.0, 1.0), c=c, mode=HIST_INT)
        x = orc.grow_levelwise(bins, nb, g=g, h=h, c=c, mode=HIST_F32_FIXED)
        assert a[0].shape[0] == depth
        for u, v, w in zip(a, q, x):
            assert np.array_equal(u, v) and np.array_equal(u, w)


def test_leafwise_adaptive_chunks_keep_every_row(orc):
    """The leaf-sized FP32 units (ORC_CHUNK_ADAPTIVE: per built node, ceil(n / T) rows
    rounded up to 256, clamped to [1024, 8192]) regroup the float sums only: with integer
    gradients every sum is exact, so any target must give the serial grower's tree (a row
    dropped or counted twice by the chunking would change a sum or a cover)."""
    bins, nb = small_plane(n=30000, seed=18)
    rng = np.random.default_rng(19)
    gq = rng.integers(-127, 128, bins.shape[0]).astype(np.int8)
    hq = rng.integers(1, 200, bins.shape[0]).astype(np.int16)
    g = gq.astype(np.float32)
    h = hq.astype(np.float32)
    c = orc_cfg(max_depth=8, min_data=5, max_leaves=40)
    a, la = orc.grow_leafwise(bins, nb, g=g, h=h, c=c, mode=HIST_F32_SERIAL)
    for target in (1, 4, 256):
        for mode in (HIST_F32_CHUNKED, HIST_F32_FIXED):
            b, lb = orc.grow_leafwise(bins, nb, g=g, h=h, c=c, mode=mode, chunk=CHUNK_ADAPTIVE | target)
            assert np.array_equal(a.view(np.uint8), b.view(np.uint8))
            assert np.array_equal(la, lb)


def leaf_groups(lor):
    """Row partition as a set of frozensets (numbering-free)."""
    return {frozenset(np.flatnonzero(lor == v).tolist()) for v in np.unique(lor)}


def test_leafwise_unbounded_equals_depthwise_partition(orc):
    """With a leaf budget above 2^depth, best-first growth splits every splittable node:
    the same row partition and leaf values as depthwise (node numbering differs)."""
    bins, nb = small_plane(n=3000, seed=12)
    rng = np.random.default_rng(13)
    g, h = orc.logloss_grad(rng.normal(0, 1, 3000).astype(np.float32),
                            (rng.random(3000) < 0.4).astype(np.float32))
    d, ld = orc.grow_depthwise(bins, nb, g=g, h=h, c=orc_cfg(max_depth=4, min_data=5))
    lw, ll = orc.grow_leafwise(bins, nb, g=g, h=h, c=orc_cfg(max_depth=4, min_data=5, max_leaves=64))
    assert leaf_groups(ld) == leaf_groups(ll)
    assert sorted(d["value"][d["feature"] < 0]) == sorted(lw["value"][lw["feature"] < 0])


def test_leafwise_budget_respected(orc):
    bins, nb = small_plane(n=3000, seed=14)
    rng = np.random.default_rng(15)
    g = rng.normal(0, 1, 3000).astype(np.float32)
    h = np.ones(3000, np.float32)
    for L in (2, 5, 9):
        t, lor = orc.grow_leafwise(bins, nb, g=g, h=h, c=orc_cfg(max_depth=8, min_data=2, max_leaves=L))
        assert int((t["feature"] < 0).sum()) == L
        assert len(np.unique(lor)) == L


def test_levelwise_is_oblivious(orc):
    bins, nb = small_plane(n=3000, seed=16)
    rng = np.random.default_rng(17)
    g = rng.normal(0, 1, 3000).astype(np.float32)
    h = np.ones(3000, np.float32)
    levels, leaves, covers, gains, lor = orc.grow_levelwise(bins, nb, g=g, h=h, c=orc_cfg(max_depth=3, min_data=2))
    D = levels.shape[0]
    # leaf index = bits of the level decisions, most significant first
    idx = np.zeros(3000, np.int64)
    for k in range(D):
        f, b, dl = levels[k]
        col = bins[:, f]
        left = np.where(col == nb[f] - 1, bool(dl), col <= b)
        idx = idx * 2 + (~left).astype(np.int64)
    assert np.array_equal(idx, lor.astype(np.int64))
    assert np.array_equal(np.bincount(idx, minlength=1 << D).astype(np.float32), covers)


def test_training_reduces_loss(orc):
    bins, nb = small_plane(n=5000, seed=18)
    y = (bins[:, 0].astype(np.float32) * 0.3 - bins[:, 1] * 0.2).astype(np.float32)
    losses = []
    for it in (1, 5, 20):
        trees, init, scores = orc.train_depthwise(bins, nb, y, 0, it, 0.1, orc_cfg(max_depth=4, min_data=5))
        losses.append(float(np.mean((scores - y) ** 2)))
        # resident scores add lr * leaf per iteration; predict sums the trees first
        assert np.allclose(scores, orc.predict(trees, bins, nb, 0.1, init), atol=1e-5, rtol=0)
    assert losses[0] > losses[1] > losses[2]


# ----------------------------------------------------------------------------- predict
def random_tree(rng, depth, F, last):
    """A random full tree of NODE_DTYPE nodes with covers that add up."""
    nodes = []

    def build(d, cover):
        i = len(nodes)
        nodes.append(None)
        if d == depth or cover < 2:
            nodes[i] = (-1, 0, 0, 0, 0, np.float32(rng.normal()), np.float32(cover), 0.0)
            return i
        f = int(rng.integers(0, F))
        b = int(rng.integers(0, last[f]))
        cl = int(rng.integers(1, cover))
        left = build(d + 1, cl)
        right = build(d + 1, cover - cl)
        nodes[i] = (f, b, int(rng.integers(0, 2)), left, right, 0.0, np.float32(cover), 1.0)
        return i

    build(0, int(rng.integers(50, 200)))
    return np.array(nodes, NODE_DTYPE)


def walk(t, row, last):
    i = 0
    while t[i]["feature"] >= 0:
        f = t[i]["feature"]
        b = row[f]
        left = bool(t[i]["default_left"]) if b == last[f] else b <= t[i]["bin"]
        i = t[i]["left"] if left else t[i]["right"]
    return t[i]["value"]


def test_predict_walk_and_rounding(orc):
    """init + lr * (sequential float sum over trees), each step rounded once."""
    rng = np.random.default_rng(19)
    F = 4
    nb = np.array([6, 7, 5, 9], np.uint16)
    last = nb - 1
    trees = [random_tree(rng, 4, F, last) for _ in range(7)]
    rows = np.stack([rng.integers(0, nb[f], 300) for f in range(F)], axis=1).astype(np.uint8)
    out = orc.predict(trees, rows, nb, 0.1, 0.25)
    for r in range(rows.shape[0]):
        acc = np.float32(0)
        for t in trees:
            acc = np.float32(acc + walk(t, rows[r], last))
        assert out[r] == np.float32(np.float32(0.25) + np.float32(np.float32(0.1) * acc))


# ----------------------------------------------------------------------------- SHAP
def exact_cond_exp(t, row, S, last):
    def rec(i):
        n = t[i]
        if n["feature"] < 0:
            return float(n["value"])
        f = n["feature"]
        if f in S:
            b = row[f]
            left = bool(n["default_left"]) if b == last[f] else b <= n["bin"]
            return rec(n["left"] if left else n["right"])
        cl, cr = float(t[n["left"]]["cover"]), float(t[n["right"]]["cover"])
        return (cl * rec(n["left"]) + cr * rec(n["right"])) / (cl + cr)
    return rec(0)


def test_shap_random_trees_brute_force(orc):
    """TreeSHAP (Algorithm 2) equals brute-force Shapley values of the cover-weighted
    conditional expectation (test_shap.cpp:92-123), and sums to the margin."""
    rng = np.random.default_rng(20)
    F = 4
    nb = np.array([5, 6, 4, 7], np.uint16)
    last = nb - 1
    for trial in range(12):
        t = random_tree(rng, 4, F, last)
        rows = np.stack([rng.integers(0, nb[f], 6) for f in range(F)], axis=1).astype(np.uint8)
        phi = orc.pred_contribs([t], rows, nb, 1.0, 0.0, 4)
        for r in range(rows.shape[0]):
            ex = np.zeros(F)
            for f in range(F):
                others = [j for j in range(F) if j != f]
                for k in range(F):
                    for S in itertools.combinations(others, k):
                        w = math.factorial(k) * math.factorial(F - k - 1) / math.factorial(F)
                        ex[f] += w * (exact_cond_exp(t, rows[r], set(S) | {f}, last)
                                      - exact_cond_exp(t, rows[r], set(S), last))
            assert np.allclose(phi[r, :F], ex, atol=1e-6, rtol=1e-6)
            # the oracle's own conditional expectation agrees with the exact walk
            s_mask = np.array([1, 0, 1, 0], np.uint8)
            ce = orc.L.orc_cond_expectation(P(t), P(np.ascontiguousarray(rows[r])), P(last), P(s_mask))
            assert ce == pytest.approx(exact_cond_exp(t, rows[r], {0, 2}, last), rel=1e-9, abs=1e-12)
        assert np.allclose(phi[:, F], exact_cond_exp(t, rows[0], set(), last), atol=1e-6)
        margin = orc.predict([t], rows, nb, 1.0, 0.0)
        assert np.allclose(phi.sum(axis=1), margin, atol=1e-5)


def test_shap_ensemble_scaling(orc):
    """phi of an ensemble = lr * sum of per-tree phi; bias = init + lr * sum E[tree]."""
    rng = np.random.default_rng(21)
    F = 3
    nb = np.array([5, 6, 4], np.uint16)
    trees = [random_tree(rng, 3, F, nb - 1) for _ in range(4)]
    rows = np.stack([rng.integers(0, nb[f], 20) for f in range(F)], axis=1).astype(np.uint8)
    ens = orc.pred_contribs(trees, rows, nb, 0.3, 0.5, 3)
    parts = sum(orc.pred_contribs([t], rows, nb, 1.0, 0.0, 3) for t in trees)
    assert np.allclose(ens[:, :F], 0.3 * parts[:, :F], atol=1e-9)
    ev = sum(orc.L.orc_tree_expected_value(P(t)) for t in trees)
    assert np.allclose(ens[:, F], 0.5 + 0.3 * ev, atol=1e-6)


# ----------------------------------------------------------------------------- AUC
def test_auc_brute_force_pairs(orc):
    rng = np.random.default_rng(22)
    from bonsai_amd import synth
    for _ in range(5):
        p = np.round(rng.random(300), 1).astype(np.float32)      # many ties
        y = (rng.random(300) < 0.4).astype(np.float32)
        pos, neg = p[y == 1], p[y == 0]
        wins = (pos[:, None] > neg[None, :]).sum() + 0.5 * (pos[:, None] == neg[None, :]).sum()
        exact = wins / (pos.size * neg.size)
        assert orc.auc(p, y) == pytest.approx(exact, abs=1e-6)
        assert synth.auc(p, y) == pytest.approx(exact, abs=1e-12)
