"""Writes reference_cases.json: the reference unit tests' known-answer cases as data.

Inputs and expected outcomes are transcribed from the reference's Catch2 tests (file and
line per case); every expected number that the reference test computes from a closed
form (split gains, SHAP weights, Shapley values, leaf weights) is computed here with
exact rational arithmetic from that closed form, never with the oracle.

    python tests/golden/make_reference_cases.py
"""
from __future__ import annotations

import itertools
import json
import math
from fractions import Fraction as Fr
from pathlib import Path

OUT = Path(__file__).resolve().parent / "reference_cases.json"
FLT_MAX = 3.4028234663852886e38
INF = "inf"


def score(g, h, l2):
    """include/bonsai/split.hpp:112-158 with lambda_l1 = 0: g^2 / (h + l2), 0 when h + l2 = 0."""
    g, h, l2 = Fr(g), Fr(h), Fr(l2)
    d = h + l2
    return g * g / d if d > 0 else Fr(0)


def cfg(min_child_hess=1.0, min_gain=0.0, l2=1.0, l1=0.0, min_data=20):
    """TreeConfig defaults (include/bonsai/config/tree_config.hpp:11-17)."""
    return {"min_child_hess": min_child_hess, "min_gain": min_gain, "l2": l2, "l1": l1, "min_data": min_data}


def hist(n, cells):
    h = [[0.0, 0.0] for _ in range(n)]
    for b, g, hh in cells:
        h[b] = [g, hh]
    return h


def split_node_cases():
    C = []
    # test_split_node.cpp:13-33 (obvious cut)
    C.append({"src": "tests/unit/test_split_node.cpp:13-33", "hists": [hist(3, [(0, -1, 1), (1, 1, 1)])],
              "cfg": cfg(l2=1, min_data=0), "n_rows": 0,
              "expect": {"valid": True, "feature": 0, "bin": 0,
                         "gain": float(score(-1, 1, 1) + score(1, 1, 1) - score(0, 2, 1))}})
    # :35-56 (best feature across two)
    C.append({"src": "tests/unit/test_split_node.cpp:35-56",
              "hists": [hist(3, [(0, -0.25, 1), (1, 0.25, 1)]), hist(3, [(0, -2, 1), (1, 2, 1)])],
              "cfg": cfg(l2=1, min_data=0), "n_rows": 0, "expect": {"valid": True, "feature": 1, "bin": 0}})
    # :58-84 (missing pulls left)
    C.append({"src": "tests/unit/test_split_node.cpp:58-84",
              "hists": [hist(3, [(0, -1, 1), (1, 1, 1), (2, -1, 1)])], "cfg": cfg(l2=1, min_data=0), "n_rows": 0,
              "expect": {"valid": True, "feature": 0, "bin": 0, "default_left": True,
                         "gain": float(score(-2, 2, 1) + score(1, 1, 1) - score(-1, 3, 1))}})
    # :86-109 (missing pulls right)
    C.append({"src": "tests/unit/test_split_node.cpp:86-109",
              "hists": [hist(3, [(0, -1, 1), (1, 1, 1), (2, 1, 1)])], "cfg": cfg(l2=1, min_data=0), "n_rows": 0,
              "expect": {"valid": True, "feature": 0, "bin": 0, "default_left": False,
                         "gain": float(score(-1, 1, 1) + score(2, 2, 1) - score(1, 3, 1))}})
    # :111-137 (no positive gain -> invalid)
    C.append({"src": "tests/unit/test_split_node.cpp:111-137", "hists": [hist(3, [(0, 0.5, 1), (1, 0.5, 1)])],
              "cfg": cfg(l2=1, min_data=0), "n_rows": 0, "expect": {"valid": False, "gain": 0.0}})
    # :153-177 (degenerate all-real-left cut skipped)
    C.append({"src": "tests/unit/test_split_node.cpp:153-177",
              "hists": [hist(3, [(0, 1, 1), (1, -1, 1), (2, 5, 0.01)])], "cfg": cfg(l2=1, min_data=0), "n_rows": 0,
              "expect": {"valid": True, "bin": 0}})
    # :179-209 (min_child_hess guard, then disabled)
    h = [hist(4, [(0, -1, 0.5), (1, 1, 2)])]
    C.append({"src": "tests/unit/test_split_node.cpp:179-202", "hists": h, "cfg": cfg(min_data=0), "n_rows": 0,
              "expect": {"valid": False}})
    C.append({"src": "tests/unit/test_split_node.cpp:204-209", "hists": h,
              "cfg": cfg(min_child_hess=0, l2=1, min_data=0), "n_rows": 0, "expect": {"valid": True, "bin": 0}})
    # :211-239 (lambda changes the cut)
    h = [hist(4, [(0, 2, 0.25), (1, 1, 4), (2, -3, 4)])]
    C.append({"src": "tests/unit/test_split_node.cpp:224-238 (lambda 0)", "hists": h,
              "cfg": cfg(min_child_hess=0, l2=0, min_data=0), "n_rows": 0, "expect": {"valid": True, "bin": 0}})
    C.append({"src": "tests/unit/test_split_node.cpp:224-238 (lambda 10)", "hists": h,
              "cfg": cfg(min_child_hess=0, l2=10, min_data=0), "n_rows": 0, "expect": {"valid": True, "bin": 1}})
    obvious = [hist(3, [(0, -1, 1), (1, 1, 1)])]
    # :259-275 / :277-294 (min_gain_to_split)
    C.append({"src": "tests/unit/test_split_node.cpp:259-275", "hists": obvious,
              "cfg": cfg(min_child_hess=0, min_gain=1.5, l2=1, min_data=0), "n_rows": 0,
              "expect": {"valid": False, "gain": 0.0}})
    C.append({"src": "tests/unit/test_split_node.cpp:277-294", "hists": obvious,
              "cfg": cfg(min_child_hess=0, min_gain=0.5, l2=1, min_data=0), "n_rows": 0,
              "expect": {"valid": True, "bin": 0, "gain": 1.0}})
    # :296-323 (min_data_in_leaf: 9 < 2*5 rejects, 100 accepts)
    C.append({"src": "tests/unit/test_split_node.cpp:296-309", "hists": obvious,
              "cfg": cfg(min_child_hess=0, l2=1, min_data=5), "n_rows": 9, "expect": {"valid": False, "gain": 0.0}})
    C.append({"src": "tests/unit/test_split_node.cpp:311-323", "hists": obvious,
              "cfg": cfg(min_child_hess=0, l2=1, min_data=5), "n_rows": 100, "expect": {"valid": True, "bin": 0}})
    return C


def split_level_cases():
    C = []
    one = hist(3, [(0, -1, 1), (1, 1, 1)])
    per = score(-1, 1, 1) + score(1, 1, 1) - score(0, 2, 1)
    C.append({"src": "tests/unit/test_split_level.cpp:26-48", "frontier": [[one]], "cfg": cfg(l2=1, min_data=0),
              "expect": {"valid": True, "feature": 0, "bin": 0, "gain": float(per)}})
    C.append({"src": "tests/unit/test_split_level.cpp:50-78", "frontier": [[one], [one]],
              "cfg": cfg(min_child_hess=0, l2=1, min_data=0),
              "expect": {"valid": True, "feature": 0, "bin": 0, "gain": float(2 * per)}})
    a = hist(4, [(0, -2, 1), (1, 0, 1), (2, 0, 1)])
    b = hist(4, [(0, 0, 1), (1, 0, 1), (2, -4, 1)])
    g = (score(-2, 2, 1) + score(0, 1, 1) - score(-2, 3, 1)) + (score(0, 2, 1) + score(-4, 1, 1) - score(-4, 3, 1))
    C.append({"src": "tests/unit/test_split_level.cpp:80-137", "frontier": [[a], [b]],
              "cfg": cfg(min_child_hess=0, l2=1, min_data=0),
              "expect": {"valid": True, "feature": 0, "bin": 1, "gain": float(g)}})
    bb = hist(3, [(0, -1, 0.5), (1, 1, 0.5)])
    C.append({"src": "tests/unit/test_split_level.cpp:139-178", "frontier": [[one], [bb]],
              "cfg": cfg(min_child_hess=0.8, l2=1, min_data=0),
              "expect": {"valid": True, "feature": 0, "bin": 0, "gain": float(per)}})
    C.append({"src": "tests/unit/test_split_level.cpp:180-203", "frontier": [[one], [one]],
              "cfg": cfg(min_child_hess=0, min_gain=3.0, l2=1, min_data=0), "expect": {"valid": False, "gain": 0.0}})
    single = hist(1, [(0, 1, 1)])
    C.append({"src": "tests/unit/test_split_level.cpp:214-235", "frontier": [[single], [single]],
              "cfg": cfg(l2=1, min_data=0), "expect": {"valid": False, "gain": 0.0}})
    return C


def leaf_weight_cases():
    """tests/unit/test_split_level.cpp:237-270 (no monotone bounds): w = -g / (h + l2), 0 on a zero denominator."""
    return [
        {"src": "tests/unit/test_split_level.cpp:249", "g": 0.0, "h": 0.0, "l1": 0.0, "l2": 0.0, "w": 0.0},
        {"src": "tests/unit/test_split_level.cpp:254", "g": 0.0, "h": 0.0, "l1": 0.5, "l2": 0.0, "w": 0.0},
        {"src": "tests/unit/test_split_level.cpp:257", "g": 3.0, "h": 0.0, "l1": 0.0, "l2": 0.0, "w": 0.0},
        {"src": "tests/unit/test_split_level.cpp:262", "g": 3.0, "h": 2.0, "l1": 0.0, "l2": 1.0, "w": -1.0},
    ]


def bin_mapper_cases():
    C = []
    C.append({"src": "tests/unit/test_bin_mapper.cpp:17-34", "column": [5, 2, 4, 1, 3], "max_bin": 255,
              "cuts": [1, 2, 3, 4, 5, FLT_MAX, INF]})
    C.append({"src": "tests/unit/test_bin_mapper.cpp:59-72", "column": [5, 4, 4, 1, 3], "max_bin": 255,
              "cuts": [1, 3, 4, 5, FLT_MAX, INF]})
    C.append({"src": "tests/unit/test_bin_mapper.cpp:74-97 (heavy value)", "column": [0.0] * 50 + list(range(1, 45)),
              "max_bin": 8, "n_cuts": 8, "transform": [[0.0, 0], [1.0, 1]]})
    C.append({"src": "tests/unit/test_bin_mapper.cpp:189-201 (single value)", "column": [3.0] * 5, "max_bin": 255,
              "cuts": [3.0, FLT_MAX, INF]})
    C.append({"src": "tests/unit/test_bin_mapper.cpp:203-212 (all NaN)", "column": ["nan"] * 5, "max_bin": 255,
              "cuts": [FLT_MAX, INF]})
    C.append({"src": "tests/unit/test_bin_mapper.cpp:248-253 (distinct path, NaN-only sentinel)",
              "column": [1.0, 2.0, 3.0], "max_bin": 255, "below_sentinel": [3.0, 6.0]})
    C.append({"src": "tests/unit/test_bin_mapper.cpp:236-246 (greedy path, capped maximum)",
              "column": [9.0] * 50 + [v / 10.0 for v in range(1, 45)], "max_bin": 8, "below_sentinel": [9.0, 18.0]})
    return C


def objective_cases():
    return [
        {"src": "tests/unit/test_objective.cpp:113-128", "objective": "logloss", "score": [0, 0, 0, 0],
         "label": [0, 1, 0, 1], "grad": [0.5, -0.5, 0.5, -0.5], "hess": [0.25] * 4, "exact": True},
        {"src": "tests/unit/test_objective.cpp:130-150", "objective": "logloss",
         "score": [math.log(3.0), -math.log(3.0)], "label": [1, 0], "grad": [-0.25, 0.25], "hess": [0.1875, 0.1875],
         "exact": False, "tol": 1e-6},
        {"src": "tests/unit/test_objective.cpp:152-170", "objective": "logloss", "score": [100, -100, 1000, -1000],
         "label": [1, 0, 0, 1], "grad": [0, 0, 1, -1], "hess": None, "exact": False, "tol": 1e-6},
    ]


def auc_cases():
    """tests/unit/test_metric.cpp:134-150."""
    lab = [0, 0, 1, 1]
    return [
        {"src": "tests/unit/test_metric.cpp:139", "p": [0.1, 0.2, 0.8, 0.9], "y": lab, "auc": 1.0},
        {"src": "tests/unit/test_metric.cpp:140", "p": [0.9, 0.8, 0.2, 0.1], "y": lab, "auc": 0.0},
        {"src": "ties (average rank)", "p": [0.5, 0.5, 0.5, 0.5], "y": lab, "auc": 0.5},
    ]


def shap_weight_table(kmax=12):
    """tests/unit/test_shap_paths.cpp:371-400: w[n][i] = 1 / (n * C(n-1, i))."""
    out = []
    for n in range(1, kmax + 1):
        for i in range(n):
            out.append(float(Fr(1, n * math.comb(n - 1, i))))
    return out


def ice_cream_tree():
    """tests/unit/test_shap_paths.cpp:421-434 and test_shap.cpp: root splits feature 1, children feature 0,
    every split at bin 0 (bin 0 left, bin 1 right, bin 2 = missing, default right), covers 100/50/50/25x4,
    leaves 0, 10, 0, 90."""
    nodes = [
        {"feature": 1, "bin": 0, "default_left": 0, "left": 1, "right": 2, "value": 0.0, "cover": 100.0},
        {"feature": 0, "bin": 0, "default_left": 0, "left": 3, "right": 4, "value": 0.0, "cover": 50.0},
        {"feature": 0, "bin": 0, "default_left": 0, "left": 5, "right": 6, "value": 0.0, "cover": 50.0},
    ]
    for v in (0.0, 10.0, 0.0, 90.0):
        nodes.append({"feature": -1, "bin": 0, "default_left": 0, "left": 0, "right": 0, "value": v, "cover": 25.0})
    return nodes


def exact_cond_expectation(nodes, row, S, last):
    def walk(i):
        n = nodes[i]
        if n["feature"] < 0:
            return Fr(n["value"])
        f = n["feature"]
        if f in S:
            b = row[f]
            left = bool(n["default_left"]) if b == last[f] else b <= n["bin"]
            return walk(n["left"] if left else n["right"])
        l, r = nodes[n["left"]], nodes[n["right"]]
        c = Fr(l["cover"]) + Fr(r["cover"])
        return (Fr(l["cover"]) * walk(n["left"]) + Fr(r["cover"]) * walk(n["right"])) / c
    return walk(0)


def exact_shapley(nodes, row, F, last):
    phi = [Fr(0)] * F
    for f in range(F):
        others = [j for j in range(F) if j != f]
        for k in range(F):
            for S in itertools.combinations(others, k):
                w = Fr(math.factorial(k) * math.factorial(F - k - 1), math.factorial(F))
                phi[f] += w * (exact_cond_expectation(nodes, row, set(S) | {f}, last)
                               - exact_cond_expectation(nodes, row, set(S), last))
    return phi


def shap_tree_cases():
    nodes = ice_cream_tree()
    last = [2, 2]
    rows = [list(r) for r in itertools.product([0, 1, 2], repeat=2)]
    phis = [[float(v) for v in exact_shapley(nodes, r, 2, last)] for r in rows]
    bias = float(exact_cond_expectation(nodes, rows[0], set(), last))
    return [{"src": "tests/unit/test_shap_paths.cpp:421-434 (ice cream); Shapley values by exact enumeration",
             "nodes": nodes, "n_bins": [3, 3], "rows": rows, "phi": phis, "bias": bias}]


def main():
    data = {
        "split_node": split_node_cases(),
        "split_level": split_level_cases(),
        "leaf_weight": leaf_weight_cases(),
        "bin_mapper": bin_mapper_cases(),
        "objective": objective_cases(),
        "auc": auc_cases(),
        "shap_weights": shap_weight_table(),
        "shap_trees": shap_tree_cases(),
    }
    OUT.write_text(json.dumps(data, indent=1) + "\n")
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
