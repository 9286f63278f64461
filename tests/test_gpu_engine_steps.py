"""The engine steps (the reference's GPULevelEngine / GPULeafEngine methods,
include/bonsai/cuda/histogram_engine.hpp:90-242) driven from the host the way the
reference's growers drive them, each tree compared with the oracle's.

- depthwise: DepthwiseGrower::grow (src/grower_impl.hpp:501-577) over LevelStep
  (src/step/level.hpp:159-366): open_level -> find_splits_many; plan_level's leaves are
  stamped; partition_level with children in slots 2k, 2k + 1; advance_level (smaller
  child built, ties to the left) or advance_layout_only at the last level; end_tree
  stamps the frontier and finalize_tree returns every row's leaf.
- levelwise: ObliviousGrower over the same plane with find_level_split.
- leafwise: LeafwiseGrower::grow (src/grower_impl.hpp:743-862) over LeafStep
  (src/step/leaf.hpp:124-267): leaf_split -> LeafRound (the smaller child takes the next
  slot), leaf_build, leaf_find of both children, leaf_stamp.

Leaf values and node totals follow the oracle's growers (its leaf_value); trees, covers,
gains and the per-row leaf are compared bit for bit in both precisions.
"""
import numpy as np
import pytest

from oracle_lib import CHUNK_ADAPTIVE, HIST_F32_FIXED, HIST_INT, NODE_DTYPE, cfg as orc_cfg
from test_gpu_depthwise import assert_same_tree, ctx_for, grad_of, tall_data
from test_gpu_growers import levelwise_view

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tall():
    return tall_data(50000, seed_cfg=9)


def leaf(v=0.0, cover=0.0):
    n = np.zeros(1, NODE_DTYPE)[0]
    n["feature"] = -1
    n["value"] = v
    n["cover"] = cover
    return n


class Totals:
    """A node's (G, H) as the oracle's leaf_value takes them."""

    def __init__(self, quantized, scales):
        self.q, self.sc = quantized, scales

    def of_node(self, info):
        rows, g, h, ig, ih = info
        return (ig * self.sc[0], ih * self.sc[1]) if self.q else (g, h)

    def of_split(self, s, side):
        if self.q:
            return (int(s["igL" if side == 0 else "igR"]) * self.sc[0],
                    int(s["ihL" if side == 0 else "ihR"]) * self.sc[1])
        return (float(np.float32(s["gL" if side == 0 else "gR"])), float(np.float32(s["hL" if side == 0 else "hR"])))


def commit_split(nodes, nid, s):
    lid = len(nodes)
    nodes += [leaf(), leaf()]
    nodes[nid]["feature"] = s["feature"]
    nodes[nid]["bin"] = s["bin"]
    nodes[nid]["default_left"] = s["default_left"]
    nodes[nid]["left"] = lid
    nodes[nid]["right"] = lid + 1
    nodes[nid]["value"] = 0.0
    nodes[nid]["gain"] = np.float32(s["gain"])
    return lid, lid + 1


def steps_depthwise(ctx, orc, c, tot):
    (g0, h0), n = ctx.begin_root()
    nodes = [leaf(cover=n)]
    cur = [{"id": 0, "depth": 0}]
    depth = 0
    while depth < c.max_depth:
        splits, info = ctx.find_splits_many([nd["depth"] for nd in cur])
        stamps, pend = [], []
        for i, nd in enumerate(cur):
            if "tot" not in nd:
                nd["tot"] = tot.of_node(info[i])
            if not splits[i]["valid"]:
                cover = nodes[nd["id"]]["cover"]
                nodes[nd["id"]] = leaf(orc.leaf_value(*nd["tot"], c), cover)
                stamps.append((i, nd["id"]))
                continue
            lid, rid = commit_split(nodes, nd["id"], splits[i])
            pend.append((i, splits[i], lid, rid))
        ctx.stamp_leaves(stamps)
        counts = ctx.partition_level([(i, 2 * k, 2 * k + 1, s["feature"], s["bin"], s["default_left"])
                                      for k, (i, s, _, _) in enumerate(pend)]) if pend else np.zeros((0, 2))
        last = depth + 1 >= c.max_depth
        nxt = []
        for k, (i, s, lid, rid) in enumerate(pend):
            nodes[lid]["cover"], nodes[rid]["cover"] = counts[k]
            L = {"id": lid, "depth": depth + 1}
            R = {"id": rid, "depth": depth + 1}
            if last:   # children are leaves: their totals are the split's child sums
                L["tot"], R["tot"] = tot.of_split(s, 0), tot.of_split(s, 1)
            nxt += [L, R]
        if last:
            ctx.advance_layout_only()
        elif pend:
            ctx.advance_level([(i, 2 * k + (0 if counts[k][0] <= counts[k][1] else 1),
                                2 * k + (1 if counts[k][0] <= counts[k][1] else 0))
                               for k, (i, s, _, _) in enumerate(pend)])
        cur = nxt
        if not cur:
            break
        depth += 1
    for i, nd in enumerate(cur):
        cover = nodes[nd["id"]]["cover"]
        nodes[nd["id"]] = leaf(orc.leaf_value(*nd["tot"], c), cover)
    ctx.stamp_leaves([(i, nd["id"]) for i, nd in enumerate(cur)])
    t = np.array(nodes, NODE_DTYPE)
    values, ids = ctx.finalize_tree(t["value"])
    return t, ids, values


@pytest.mark.parametrize("quantized", [True, False])
@pytest.mark.parametrize("depth", [1, 4, 7])
def test_depthwise_steps_match_the_oracle(hip, orc, tall, quantized, depth):
    bins, nb, X, y = tall
    g, h = grad_of(orc, y, seed=50 + depth)
    c = orc_cfg(max_depth=depth)
    ctx = ctx_for(hip, bins, nb, quantized=quantized, max_depth=depth)
    if quantized:
        gq, hq, sc = ctx.set_gradients(g, h)
        ref, rlor = orc.grow_depthwise(bins, nb, gq=gq, hq=hq, scales=sc, c=c, mode=HIST_INT)
    else:
        ctx.set_gradients(g, h)
        sc = (0.0, 0.0)
        ref, rlor = orc.grow_depthwise(bins, nb, g=g, h=h, c=c, mode=HIST_F32_FIXED)
    tree, ids, values = steps_depthwise(ctx, orc, c, Totals(quantized, sc))
    coarse, clor = ctx.grow_tree()
    ctx.close()
    assert_same_tree(tree, ref)
    assert np.array_equal(ids, rlor)
    assert np.array_equal(values.view(np.uint32), ref["value"][rlor].view(np.uint32))
    assert_same_tree(coarse, tree)            # the device-resident grower builds the same tree


def test_depthwise_steps_on_a_row_subset(hip, orc, tall):
    """begin_root over the sampler's rows: the tree is grown on them, every row is routed."""
    bins, nb, X, y = tall
    g, h = grad_of(orc, y, seed=61)
    rows = np.flatnonzero(np.random.default_rng(2).random(y.size) < 0.4).astype(np.uint32)
    c = orc_cfg(max_depth=5)
    ctx = ctx_for(hip, bins, nb, quantized=True, max_depth=5)
    gq, hq, sc = ctx.set_gradients(g, h)
    ctx.set_row_subset(rows)
    tree, ids, _ = steps_depthwise(ctx, orc, c, Totals(True, sc))
    ctx.close()
    ref, rlor = orc.grow_depthwise(np.ascontiguousarray(bins[rows]), nb, gq=gq[rows], hq=hq[rows], scales=sc, c=c,
                                   mode=HIST_INT)
    assert_same_tree(tree, ref)
    assert np.array_equal(ids[rows], rlor)


def steps_levelwise(ctx, orc, c, tot, per_node=False):
    ctx.begin_root()
    cur = [0]
    levels, gains = [], []
    tots = None
    for depth in range(c.max_depth):
        splits, info = ctx.find_level_split([depth] * len(cur))
        if not splits[0]["valid"]:
            tots = [tot.of_node(x) for x in info]
            break
        s0 = splits[0]
        levels.append((int(s0["feature"]), int(s0["bin"]), int(s0["default_left"])))
        gains.append(np.float32(s0["gain"]))
        last = depth + 1 >= c.max_depth
        counts = ctx.partition_level([(i, 2 * i, 2 * i + 1, s0["feature"], s0["bin"], s0["default_left"])
                                      for i in range(len(cur))])
        if last:
            tots = [t for i in range(len(cur)) for t in (tot.of_split(splits[i], 0), tot.of_split(splits[i], 1))]
            ctx.advance_layout_only()
        else:
            ctx.advance_level([(i, 2 * i + (0 if counts[i][0] <= counts[i][1] else 1),
                                2 * i + (1 if counts[i][0] <= counts[i][1] else 0)) for i in range(len(cur))],
                              per_node=per_node)
        cur = list(range(2 * len(cur)))
    if tots is None:   # the loop ran out of depth after a find: totals of the frontier
        _, info = ctx.find_splits_many([c.max_depth] * len(cur))
        tots = [tot.of_node(x) for x in info]
    leaves = np.array([orc.leaf_value(g, h, c) for g, h in tots], np.float32)
    ctx.stamp_leaves([(i, i) for i in range(len(cur))])
    _, ids = ctx.finalize_tree(leaves)
    return np.array(levels, np.int64).reshape(-1, 3), leaves, ids


@pytest.mark.parametrize("per_node", [False, True])
@pytest.mark.parametrize("quantized", [True, False])
def test_levelwise_steps_match_the_oracle(hip, orc, tall, quantized, per_node):
    """per_node: one build and one subtraction launch per node (the reference's cadence)."""
    bins, nb, X, y = tall
    g, h = grad_of(orc, y, seed=71)
    depth = 5
    c = orc_cfg(max_depth=depth)
    ctx = ctx_for(hip, bins, nb, grower="hip_levelwise", quantized=quantized, max_depth=depth)
    if quantized:
        gq, hq, sc = ctx.set_gradients(g, h)
        ref = orc.grow_levelwise(bins, nb, gq=gq, hq=hq, scales=sc, c=c, mode=HIST_INT)
    else:
        ctx.set_gradients(g, h)
        sc = (0.0, 0.0)
        ref = orc.grow_levelwise(bins, nb, g=g, h=h, c=c, mode=HIST_F32_FIXED)
    levels, leaves, ids = steps_levelwise(ctx, orc, c, Totals(quantized, sc), per_node)
    ctx.close()
    assert np.array_equal(levels, ref[0].astype(np.int64))
    assert np.array_equal(leaves.view(np.uint32), ref[1].view(np.uint32))
    assert np.array_equal(ids, ref[4])


def steps_leafwise(ctx, orc, c, tot):
    (g0, h0), n = ctx.begin_root(leaf_plane=True)
    nodes = [leaf(cover=n)]
    root = {"id": 0, "depth": 0, "slot": 0}
    s, info = ctx.leaf_find([0], [0])
    root["tot"], root["split"] = tot.of_node(info[0]), s[0]
    heap, pending = ([root], []) if s[0]["valid"] and c.max_depth > 0 else ([], [root])
    live = 1
    while heap and (c.max_leaves == 0 or live < c.max_leaves):
        best = 0
        for i in range(1, len(heap)):
            a, b = heap[i], heap[best]
            if a["split"]["gain"] > b["split"]["gain"] or (a["split"]["gain"] == b["split"]["gain"]
                                                           and a["id"] < b["id"]):
                best = i
        cd = heap.pop(best)
        sp = cd["split"]
        lid, rid = commit_split(nodes, cd["id"], sp)
        ls, rs, nl, nr = ctx.leaf_split(cd["slot"], int(sp["feature"]), int(sp["bin"]), int(sp["default_left"]))
        if nl == 0 or nr == 0:          # demoted back to a leaf (commit_pop)
            cover = nodes[cd["id"]]["cover"]
            nodes[cd["id"]] = leaf(orc.leaf_value(*cd["tot"], c), cover)
            pending.append(cd)
            continue
        L = {"id": lid, "depth": cd["depth"] + 1, "slot": ls}
        R = {"id": rid, "depth": cd["depth"] + 1, "slot": rs}
        nodes[lid]["cover"], nodes[rid]["cover"] = nl, nr
        live += 1
        if L["depth"] >= c.max_depth:
            L["tot"], R["tot"] = tot.of_split(sp, 0), tot.of_split(sp, 1)
            L["split"] = R["split"] = None
            pending += [L, R]
            continue
        small, large = (ls, rs) if nl <= nr else (rs, ls)
        ctx.leaf_build(small, large)
        s2, info2 = ctx.leaf_find([L["depth"], R["depth"]], [ls, rs])
        for nd, sx, inf in ((L, s2[0], info2[0]), (R, s2[1], info2[1])):
            nd["tot"], nd["split"] = tot.of_node(inf), sx
            (heap if sx["valid"] else pending).append(nd)
    for nd in heap + pending:
        cover = nodes[nd["id"]]["cover"]
        nodes[nd["id"]] = leaf(orc.leaf_value(*nd["tot"], c), cover)
    ctx.stamp_leaves([(nd["slot"], nd["id"]) for nd in heap + pending])
    t = np.array(nodes, NODE_DTYPE)
    values, ids = ctx.finalize_tree(t["value"])
    return t, ids


@pytest.mark.parametrize("quantized", [True, False])
@pytest.mark.parametrize("leaves,depth", [(2, 12), (31, 12), (63, 6)])
def test_leafwise_steps_match_the_oracle(hip, orc, tall, quantized, leaves, depth):
    bins, nb, X, y = tall
    g, h = grad_of(orc, y, seed=80 + leaves)
    c = orc_cfg(max_depth=depth, max_leaves=leaves)
    ctx = ctx_for(hip, bins, nb, grower="hip_leafwise", quantized=quantized, max_depth=depth, max_leaves=leaves)
    if quantized:
        gq, hq, sc = ctx.set_gradients(g, h)
        ref, rlor = orc.grow_leafwise(bins, nb, gq=gq, hq=hq, scales=sc, c=c, mode=HIST_INT)
    else:
        ctx.set_gradients(g, h)
        sc = (0.0, 0.0)
        # the steps build with 8,192-row units (the fixed-chunk order)
        ref, rlor = orc.grow_leafwise(bins, nb, g=g, h=h, c=c, mode=HIST_F32_FIXED)
    tree, ids = steps_leafwise(ctx, orc, c, Totals(quantized, sc))
    ctx.close()
    assert_same_tree(tree, ref)
    assert np.array_equal(ids, rlor)


def test_steps_refuse_what_the_host_grower_owns(hip, orc, tall):
    bins, nb, X, y = tall
    ctx = ctx_for(hip, bins, nb, quantized=True, max_depth=4)
    g, h = grad_of(orc, y, seed=90)
    ctx.set_gradients(g, h)
    with pytest.raises(hip.ConfigError):
        ctx.partition_level([(0, 0, 1, 0, 3, 1)])          # no tree begun
    ctx.begin_root()
    with pytest.raises(hip.ConfigError):
        ctx.partition_level([(5, 0, 1, 0, 3, 1)])          # no such slot
    with pytest.raises(hip.ConfigError):
        ctx.begin_root(leaf_plane=True)                    # a depthwise fit has no leaf plane
    ctx.set_monotone([1])
    with pytest.raises(hip.ConfigError):
        ctx.begin_root()
    ctx.close()
