"""The oracle's interaction constraints (tree.interaction_constraints): a node may split
on the union of the groups that contain every distinct feature on its path, plus that
feature when the path has one, the root on any (allowed_features,
src/grower_impl.hpp:111-137; propagate_interaction_state, :143-161; the finder's skip,
src/split.cpp:111-114).

Pinned by the reference's grower case (tests/unit/test_grower.cpp:368-417: eight rows,
groups "0,1" and "2", no root-to-leaf path mixes feature 2 with 0 or 1), by the
invariant the rule implies for every node of every constrained tree (its path's
feature set is one feature or lies inside one group) on larger data where the free
tree breaks it, by the config-string parsing, and by no groups growing the free tree
bit for bit.
"""
import numpy as np
import pytest

from bonsai_amd import parse_interaction_groups
from oracle_lib import HIST_F32_SERIAL, HIST_INT, cfg as orc_cfg


def path_sets(nodes):
    """Feature set of the path above every node, root first."""
    out = {}
    stack = [(0, frozenset())]
    while stack:
        i, used = stack.pop()
        out[i] = used
        if nodes["feature"][i] >= 0:
            u = used | {int(nodes["feature"][i])}
            stack.append((int(nodes["left"][i]), u))
            stack.append((int(nodes["right"][i]), u))
    return out


def violations(nodes, groups):
    bad = 0
    for used in path_sets(nodes).values():
        if len(used) > 1 and not any(used <= set(g) for g in groups):
            bad += 1
    return bad


def with_groups(orc, groups, fn, *a, **kw):
    orc.set_interaction(groups)
    try:
        return fn(*a, **kw)
    finally:
        orc.set_interaction(None)


def test_parse_like_the_config():
    assert parse_interaction_groups(["0,1", "2"]) == [[0, 1], [2]]
    assert parse_interaction_groups(["3+1", "", "5,,2"]) == [[1, 3], [2, 5]]
    assert parse_interaction_groups([[4, 0], []]) == [[0, 4]]


def test_reference_case_keeps_groups_on_separate_paths(orc):
    cols = [[0.0, 0.1, 0.2, 0.3, 1.0, 1.1, 1.2, 1.3],
            [0.0, 1.0, 0.1, 1.1, 0.2, 1.2, 0.3, 1.3],
            [0.5, 0.0, 1.0, 0.6, 0.1, 1.1, 0.7, 1.2]]
    cuts = [orc.fit_cuts(np.array(c, np.float32), 255) for c in cols]
    bins = np.stack([orc.transform(cuts[j], np.array(cols[j], np.float32)) for j in range(3)], axis=1)
    nb = np.array([c.size for c in cuts], np.uint16)
    g = np.array([-3.0, 1.0, -2.0, 2.0, -1.0, 3.0, -2.5, 1.5], np.float32)
    h = np.ones(8, np.float32)
    c = orc_cfg(min_child_hess=0.0, l2=1.0, max_depth=3, min_data=0)
    groups = parse_interaction_groups(["0,1", "2"])
    tree, _ = with_groups(orc, groups, orc.grow_depthwise, bins, nb, g=g, h=h, c=c, mode=HIST_F32_SERIAL)
    assert (tree["feature"] >= 0).sum() >= 1
    for used in path_sets(tree).values():
        assert not (2 in used and (0 in used or 1 in used))


def mixed(orc, n=8000, F=6, seed=5):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(n, F)).astype(np.float32)
    y = (X[:, 0] * X[:, 3] + X[:, 1] * X[:, 4] + np.sin(2 * X[:, 2]) + X[:, 5]
         + 0.1 * rng.normal(size=n)).astype(np.float32)
    cuts = [orc.fit_cuts(X[:, j], 63) for j in range(F)]
    bins = np.stack([orc.transform(cuts[j], X[:, j]) for j in range(F)], axis=1)
    nb = np.array([c.size for c in cuts], np.uint16)
    g, h = orc.mse_grad(np.zeros(n, np.float32), y)
    return bins, nb, g, h


@pytest.mark.parametrize("grower", ["grow_depthwise", "grow_leafwise"])
@pytest.mark.parametrize("quantized", [True, False])
def test_every_path_stays_inside_a_group(orc, grower, quantized):
    bins, nb, g, h = mixed(orc)
    c = orc_cfg(max_depth=6, max_leaves=40, min_data=10)
    fn = getattr(orc, grower)
    kw = dict(c=c)
    if quantized:
        gq, hq, sc = orc.quantize(g, h, 13, 0)
        kw.update(gq=gq, hq=hq, scales=sc, mode=HIST_INT)
    else:
        kw.update(g=g, h=h, mode=HIST_F32_SERIAL)
    groups = [[0, 1, 2], [3, 4], [2, 5]]
    free, _ = fn(bins, nb, **kw)
    assert violations(free, groups) > 0
    tree, _ = with_groups(orc, groups, fn, bins, nb, **kw)
    assert violations(tree, groups) == 0
    assert len({int(f) for f in tree["feature"] if f >= 0}) >= 2
    none, _ = with_groups(orc, [], fn, bins, nb, **kw)
    assert none.tobytes() == free.tobytes()
