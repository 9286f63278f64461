"""The oracle's monotone constraints (tree.monotone_constraints): the finder skips a
candidate whose bounded child weights break its feature's direction (src/split.cpp:
143-153), children inherit the parent's leaf-weight box fenced at the midpoint of the
child weights (propagate_monotone_bounds, src/grower_impl.hpp:47-75), and leaf values
are clamped to their box (bounded_leaf_weight, include/bonsai/split.hpp:141-150).

Pinned by the reference's own grower cases (tests/unit/test_grower.cpp:286-363: nine
rows of one feature whose unconstrained tree is non-monotone; +1 depthwise and -1
leafwise give monotone prediction curves), by the property every constrained tree must
have on larger data (walking a row through every bin of a constrained feature never
moves its prediction against the direction), and by all-free directions growing the
unconstrained tree bit for bit.
"""
import numpy as np
import pytest

from oracle_lib import HIST_F32_SERIAL, HIST_INT, cfg as orc_cfg


def leaf_value(nodes, row_bins, nb):
    i = 0
    while nodes["feature"][i] >= 0:
        f = nodes["feature"][i]
        b = int(row_bins[f])
        if b == int(nb[f]) - 1:
            left = nodes["default_left"][i] != 0
        else:
            left = b <= nodes["bin"][i]
        i = nodes["left"][i] if left else nodes["right"][i]
    return float(nodes["value"][i])


def nine_rows(orc):
    x = np.array([0.0, 0.1, 0.2, 1.0, 1.1, 1.2, 2.0, 2.1, 2.2], np.float32)
    cuts = orc.fit_cuts(x, 255)
    bins = orc.transform(cuts, x).reshape(-1, 1)
    nb = np.array([cuts.size], np.uint16)
    return bins, nb


def curve(nodes, bins, nb):
    # one row from each third, as the reference evaluates x = 0.0, 1.05, 2.1
    return [leaf_value(nodes, bins[k], nb) for k in (0, 4, 7)]


def with_monotone(orc, dirs, fn, *a, **kw):
    orc.set_monotone(dirs)
    try:
        return fn(*a, **kw)
    finally:
        orc.set_monotone(None)


def test_depthwise_plus_one_gives_a_non_decreasing_curve(orc):
    bins, nb = nine_rows(orc)
    g = np.array([-1, -1, -1, 2, 2, 2, -2, -2, -2], np.float32)
    h = np.ones(9, np.float32)
    c = orc_cfg(min_child_hess=0.0, l2=1.0, max_depth=4, min_data=0)
    free, _ = orc.grow_depthwise(bins, nb, g=g, h=h, c=c, mode=HIST_F32_SERIAL)
    fc = curve(free, bins, nb)
    assert fc[1] < fc[0] or fc[2] < fc[1]  # the unconstrained tree is non-monotone
    mono, _ = with_monotone(orc, [1], orc.grow_depthwise, bins, nb, g=g, h=h, c=c, mode=HIST_F32_SERIAL)
    mc = curve(mono, bins, nb)
    assert mc[0] <= mc[1] <= mc[2]


def test_leafwise_minus_one_gives_a_non_increasing_curve(orc):
    bins, nb = nine_rows(orc)
    g = np.array([1, 1, 1, -2, -2, -2, 2, 2, 2], np.float32)
    h = np.ones(9, np.float32)
    c = orc_cfg(min_child_hess=0.0, l2=1.0, max_depth=4, min_data=0, max_leaves=8)
    free, _ = orc.grow_leafwise(bins, nb, g=g, h=h, c=c, mode=HIST_F32_SERIAL)
    fc = curve(free, bins, nb)
    assert fc[1] > fc[0] or fc[2] > fc[1]
    mono, _ = with_monotone(orc, [-1], orc.grow_leafwise, bins, nb, g=g, h=h, c=c, mode=HIST_F32_SERIAL)
    mc = curve(mono, bins, nb)
    assert mc[0] >= mc[1] >= mc[2]


def wavy(orc, n=8000, F=6, seed=3):
    """Targets non-monotone in features 0 and 2, so constraints bite."""
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(n, F)).astype(np.float32)
    y = (np.sin(2.5 * X[:, 0]) + X[:, 1] - np.cos(2.0 * X[:, 2]) + 0.1 * rng.normal(size=n)).astype(np.float32)
    cuts = [orc.fit_cuts(X[:, j], 63) for j in range(F)]
    bins = np.stack([orc.transform(cuts[j], X[:, j]) for j in range(F)], axis=1)
    nb = np.array([c.size for c in cuts], np.uint16)
    g, h = orc.mse_grad(np.zeros(n, np.float32), y)
    return bins, nb, g, h


def monotone_violations(nodes, bins, nb, f, direction, n_rows=200):
    bad = 0
    for r in range(n_rows):
        row = bins[r].copy()
        prev = None
        for b in range(int(nb[f]) - 1):  # real bins in order (the missing bin is not ordered)
            row[f] = b
            v = leaf_value(nodes, row, nb)
            if prev is not None and direction * (v - prev) < 0:
                bad += 1
            prev = v
    return bad


@pytest.mark.parametrize("grower", ["grow_depthwise", "grow_leafwise"])
@pytest.mark.parametrize("quantized", [True, False])
def test_constrained_trees_are_monotone(orc, grower, quantized):
    bins, nb, g, h = wavy(orc)
    c = orc_cfg(max_depth=7, max_leaves=48, min_data=10)
    fn = getattr(orc, grower)
    kw = dict(c=c)
    if quantized:
        gq, hq, sc = orc.quantize(g, h, 11, 0)
        kw.update(gq=gq, hq=hq, scales=sc, mode=HIST_INT)
    else:
        kw.update(g=g, h=h, mode=HIST_F32_SERIAL)
    free, _ = fn(bins, nb, **kw)
    assert monotone_violations(free, bins, nb, 0, 1) > 0 and monotone_violations(free, bins, nb, 2, -1) > 0
    dirs = [1, 0, -1, 0, 0, 0]
    mono, lor = with_monotone(orc, dirs, fn, bins, nb, **kw)
    assert (mono["feature"] == 0).any() and (mono["feature"] == 2).any()  # still used, but monotone
    assert monotone_violations(mono, bins, nb, 0, 1) == 0
    assert monotone_violations(mono, bins, nb, 2, -1) == 0
    # all-free directions are the unconstrained grower
    zero, _ = with_monotone(orc, [0] * 6, fn, bins, nb, **kw)
    assert zero.tobytes() == free.tobytes()
