"""The oracle against the reference's own known-answer cases (tests/golden/reference_cases.json).

Each case names the reference test it comes from.  Expected gains, weights and Shapley
values were computed from the closed forms with exact rational arithmetic by
tests/golden/make_reference_cases.py, not by the oracle.
"""
import ctypes as C
import json
import math
from pathlib import Path

import numpy as np
import pytest

from oracle_lib import NODE_DTYPE, P, TreeCfg

CASES = json.loads((Path(__file__).resolve().parent / "golden" / "reference_cases.json").read_text())


def tcfg(c, max_depth=6):
    return TreeCfg(c["min_child_hess"], c["min_gain"], c["l2"], c["l1"], max_depth, c["min_data"], 31, 0)


def as_hist(feats):
    """[feature][bin][g,h] padded to the widest feature; n_bins per feature."""
    nb = np.array([len(f) for f in feats], np.uint16)
    SB = int(nb.max())
    h = np.zeros((len(feats), SB, 2), np.float32)
    for j, f in enumerate(feats):
        h[j, : len(f)] = np.asarray(f, np.float32)
    return h, nb


def node_totals(h, nb):
    """Node (G, H): double sum of the first feature's cells, rounded to float
    (include/bonsai/histogram.hpp:76-87)."""
    g = float(np.float32(np.sum(h[0, : nb[0], 0].astype(np.float64))))
    hh = float(np.float32(np.sum(h[0, : nb[0], 1].astype(np.float64))))
    return g, hh


def check_split(s, e):
    assert bool(s.valid) == e["valid"]
    if "feature" in e:
        assert s.feature == e["feature"]
    if "bin" in e:
        assert s.bin == e["bin"]
    if "default_left" in e:
        assert bool(s.default_left) == e["default_left"]
    if "gain" in e:
        assert s.gain == pytest.approx(e["gain"], rel=1e-12, abs=1e-12)


@pytest.mark.parametrize("case", CASES["split_node"], ids=lambda c: c["src"])
def test_split_node_f32(orc, case):
    h, nb = as_hist(case["hists"])
    g, hh = node_totals(h, nb)
    s = orc.find_node_f32(h, nb, g, hh, case["n_rows"], tcfg(case["cfg"]))
    check_split(s, case["expect"])


@pytest.mark.parametrize("case", CASES["split_node"], ids=lambda c: c["src"])
def test_split_node_int_same_decisions(orc, case):
    """Fixed-point cells (quantum 1/400 holds every case's decimal values) give the same decisions."""
    h, nb = as_hist(case["hists"])
    q = np.rint(h.astype(np.float64) * 400).astype(np.int64)
    assert np.allclose(q / 400.0, h.astype(np.float64), rtol=1e-7, atol=0)
    tg, th = int(q[0, : nb[0], 0].sum()), int(q[0, : nb[0], 1].sum())
    s = orc.find_node_int(q, nb, tg, th, 1 / 400, 1 / 400, case["n_rows"], tcfg(case["cfg"]))
    check_split(s, case["expect"])


@pytest.mark.parametrize("case", CASES["split_level"], ids=lambda c: c["src"])
def test_split_level_f32(orc, case):
    nodes = [as_hist(n) for n in case["frontier"]]
    nb = nodes[0][1]
    H = np.stack([h for h, _ in nodes])
    T = np.array([node_totals(h, nb) for h, _ in nodes], np.float32)
    s = orc.find_level_f32(H, T, nb, tcfg(case["cfg"]))
    check_split(s, case["expect"])


@pytest.mark.parametrize("case", CASES["split_level"], ids=lambda c: c["src"])
def test_split_level_int_same_decisions(orc, case):
    """The integer level finder on the same cases in fixed point (quantum 1/400 holds
    every case's decimal values exactly): the same decision and gain."""
    nodes = [as_hist(n) for n in case["frontier"]]
    nb = nodes[0][1]
    H = np.stack([h for h, _ in nodes])
    q = np.rint(H.astype(np.float64) * 400).astype(np.int64)
    assert np.allclose(q / 400.0, H.astype(np.float64), rtol=1e-7, atol=0)
    T = np.array([[int(n[0, : nb[0], 0].sum()), int(n[0, : nb[0], 1].sum())] for n in q], np.int64)
    s = orc.find_level_int(q, T, nb, 1 / 400, 1 / 400, tcfg(case["cfg"]))
    check_split(s, case["expect"])


def test_level_single_parent_equals_node_finder(orc):
    """test_split_level.cpp:26-48 on random histograms: a one-node frontier is the node finder."""
    rng = np.random.default_rng(0)
    for _ in range(50):
        F, SB = 3, 9
        h = np.zeros((F, SB, 2), np.float32)
        h[:, :, 0] = rng.normal(0, 1, (F, SB))
        h[:, :, 1] = rng.uniform(0.5, 2, (F, SB))
        # every feature holds the same rows: same totals
        h[1:, :, :] *= (h[0].sum(axis=0) / h[1:].sum(axis=1))[:, None, :]
        nb = np.full(F, SB, np.uint16)
        g, hh = node_totals(h, nb)
        c = TreeCfg(0.0, 0.0, 1.0, 0.0, 6, 0, 31, 0)
        a = orc.find_node_f32(h, nb, g, hh, 0, c)
        b = orc.find_level_f32(h[None], np.array([[g, hh]], np.float32), nb, c)
        assert (a.valid, a.feature, a.bin, a.default_left) == (b.valid, b.feature, b.bin, b.default_left)
        if a.valid:
            assert a.gain == pytest.approx(b.gain, rel=1e-6)


def test_level_int_single_parent_equals_node_int_finder(orc):
    """The integer twins: a one-node frontier of the integer level finder is the integer
    node finder (decision and gain), on random fixed-point histograms."""
    rng = np.random.default_rng(1)
    for _ in range(50):
        F, SB = 4, 11
        base = np.zeros((SB, 2), np.int64)
        base[:, 0] = rng.integers(-500, 500, SB)
        base[:, 1] = rng.integers(50, 900, SB)
        # every feature holds the same rows: permute one node's cells across bins
        q = np.stack([base[rng.permutation(SB)] for _ in range(F)])
        nb = np.full(F, SB, np.uint16)
        tg, th = int(base[:, 0].sum()), int(base[:, 1].sum())
        c = TreeCfg(0.0, 0.0, 1.0, 0.0, 6, 0, 31, 0)
        a = orc.find_node_int(q, nb, tg, th, 0.01, 0.002, 0, c)
        b = orc.find_level_int(q[None], np.array([[tg, th]], np.int64), nb, 0.01, 0.002, c)
        assert (a.valid, a.feature, a.bin, a.default_left) == (b.valid, b.feature, b.bin, b.default_left)
        if a.valid:
            assert a.gain == pytest.approx(b.gain, rel=1e-12)


@pytest.mark.parametrize("case", CASES["leaf_weight"], ids=lambda c: c["src"])
def test_leaf_weight(orc, case):
    c = TreeCfg(0.0, 0.0, case["l2"], case["l1"], 6, 0, 31, 0)
    assert orc.leaf_value(case["g"], case["h"], c) == case["w"]


def col_of(case):
    return np.array([np.nan if v == "nan" else v for v in case["column"]], np.float32)


def expected_cuts(case):
    return np.array([np.inf if v == "inf" else v for v in case["cuts"]], np.float32)


@pytest.mark.parametrize("case", CASES["bin_mapper"], ids=lambda c: c["src"])
def test_bin_mapper(orc, case):
    from bonsai_amd import binning
    col = col_of(case)
    cuts = orc.fit_cuts(col, case["max_bin"])
    host = binning.create_cuts(col[~np.isnan(col)], case["max_bin"])
    assert np.array_equal(cuts, host), "host mirror and oracle disagree"
    assert np.all(np.diff(cuts.astype(np.float64)) > 0)
    assert cuts[-1] == np.inf and cuts.size <= case["max_bin"]
    if "cuts" in case:
        assert np.array_equal(cuts, expected_cuts(case))
    if "n_cuts" in case:
        assert cuts.size == case["n_cuts"]
    for x, b in case.get("transform", []):
        assert orc.transform(cuts, np.array([x], np.float32))[0] == b
    sentinel = cuts.size - 1
    for x in case.get("below_sentinel", []):
        assert orc.transform(cuts, np.array([x], np.float32))[0] < sentinel
    assert orc.transform(cuts, np.array([np.nan], np.float32))[0] == sentinel


@pytest.mark.parametrize("case", CASES["objective"], ids=lambda c: c["src"])
def test_objective(orc, case):
    s = np.array(case["score"], np.float32)
    y = np.array(case["label"], np.float32)
    g, h = orc.logloss_grad(s, y)
    assert np.all(np.isfinite(g)) and np.all(np.isfinite(h))
    if case["exact"]:
        assert np.array_equal(g, np.array(case["grad"], np.float32))
        assert np.array_equal(h, np.array(case["hess"], np.float32))
    else:
        assert np.allclose(g, case["grad"], atol=case["tol"], rtol=0)
        if case["hess"] is not None:
            assert np.allclose(h, case["hess"], atol=case["tol"], rtol=0)


@pytest.mark.parametrize("case", CASES["auc"], ids=lambda c: c["src"])
def test_auc(orc, case):
    from bonsai_amd import synth
    p = np.array(case["p"], np.float32)
    y = np.array(case["y"], np.float32)
    assert orc.auc(p, y) == pytest.approx(case["auc"], abs=1e-7)
    assert synth.auc(p, y) == pytest.approx(case["auc"], abs=1e-12)


def test_shap_weight_table(orc):
    w = np.zeros(78, np.float64)
    orc.L.orc_shap_path_weights(C.c_uint32(12), P(w))
    assert np.array_equal(w, np.array(CASES["shap_weights"]))
    # Shapley normalisation: each row sums to one over its C(n-1, i) subsets
    base = 0
    for n in range(1, 13):
        assert sum(math.comb(n - 1, i) * w[base + i] for i in range(n)) == pytest.approx(1.0, abs=1e-12)
        base += n


def golden_tree(case):
    t = np.zeros(len(case["nodes"]), NODE_DTYPE)
    for i, n in enumerate(case["nodes"]):
        t[i] = (n["feature"], n["bin"], n["default_left"], n["left"], n["right"], n["value"], n["cover"], 0.0)
    return t


@pytest.mark.parametrize("case", CASES["shap_trees"], ids=lambda c: c["src"])
def test_shap_golden_tree(orc, case):
    t = golden_tree(case)
    rows = np.array(case["rows"], np.uint8)
    nb = np.array(case["n_bins"], np.uint16)
    phi = orc.pred_contribs([t], rows, nb, 1.0, 0.0, 4)
    assert np.allclose(phi[:, :-1], np.array(case["phi"]), atol=1e-9, rtol=0)
    assert np.allclose(phi[:, -1], case["bias"], atol=1e-9)
    # additivity: sum(phi) = margin
    margin = orc.predict([t], rows, nb, 1.0, 0.0)
    assert np.allclose(phi.sum(axis=1), margin, atol=1e-9)
