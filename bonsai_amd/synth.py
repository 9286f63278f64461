"""Synthetic workloads of the benchmark configurations (seeded, numpy PCG64).

Seed sequence per workload: SeedSequence([42, config_id, n_rows, n_features, stream]),
stream 0 = training rows, 1 = held-out rows.  Columns are drawn one at a time in
float32 so the 16M x 28 matrix costs 1.8 GB.

tall (configs 0-2): features 0-9 N(0,1), 10-17 U(-1,1), 18-23 lognormal(0,1), 24-27
integers in [0,16); 5% of the cells of features 18-27 are NaN.  A NaN operand makes
its label term zero.
  regression target: sin x0 + x1*x2 + 0.5|x10| + 0.1 log x18 + N(0, 0.1^2)
  binary label: Bernoulli(sigmoid(2 (sin x0 + x1*x2 - 0.5 x10^2 + 0.3 x24/16)))
wide (config 3): 2,000 features of 64 bin values: 90% sparse (0 w.p. 0.95, else
U{1..63}), 10% dense U{0..63}; target = sum of 50 random depth-3 interactions +
N(0, 0.1^2).
"""
from __future__ import annotations

import numpy as np


def _rng(config_id: int, n: int, f: int, stream: int) -> np.random.Generator:
    return np.random.default_rng(np.random.SeedSequence([42, config_id, n, f, stream]))


def tall_features(n: int, config_id: int = 1, stream: int = 0) -> np.ndarray:
    f = 28
    rng = _rng(config_id, n, f, stream)
    X = np.empty((n, f), np.float32)
    for j in range(10):
        X[:, j] = rng.standard_normal(n, dtype=np.float32)
    for j in range(10, 18):
        X[:, j] = rng.random(n, dtype=np.float32) * np.float32(2) - np.float32(1)
    for j in range(18, 24):
        X[:, j] = np.exp(rng.standard_normal(n, dtype=np.float32))
    for j in range(24, 28):
        X[:, j] = rng.integers(0, 16, n).astype(np.float32)
    for j in range(18, 28):
        X[rng.random(n, dtype=np.float32) < np.float32(0.05), j] = np.nan
    return X


def _nz(v: np.ndarray) -> np.ndarray:
    return np.where(np.isnan(v), np.float32(0), v)


def regression_target(X: np.ndarray, config_id: int = 0, stream: int = 0) -> np.ndarray:
    n = X.shape[0]
    rng = _rng(config_id, n, X.shape[1], 100 + stream)
    with np.errstate(invalid="ignore", divide="ignore"):
        t18 = _nz(np.float32(0.1) * np.log(X[:, 18]))
    y = (np.sin(X[:, 0]) + X[:, 1] * X[:, 2] + np.float32(0.5) * np.abs(X[:, 10]) + t18
         + np.float32(0.1) * rng.standard_normal(n, dtype=np.float32))
    return y.astype(np.float32)


def binary_label(X: np.ndarray, config_id: int = 1, stream: int = 0) -> np.ndarray:
    n = X.shape[0]
    rng = _rng(config_id, n, X.shape[1], 100 + stream)
    z = np.float32(2) * (np.sin(X[:, 0]) + X[:, 1] * X[:, 2] - np.float32(0.5) * X[:, 10] ** 2
                         + np.float32(0.3) * _nz(X[:, 24]) / np.float32(16))
    p = np.float32(1) / (np.float32(1) + np.exp(-z))
    return (rng.random(n, dtype=np.float32) < p).astype(np.float32)


def wide_bins(n: int, f: int = 2000, config_id: int = 3, stream: int = 0):
    """Config 3: raw values 0..63 as float32 (binned by the caller) and the target."""
    rng = _rng(config_id, n, f, stream)
    n_sparse = int(round(0.9 * f))
    X = np.empty((n, f), np.float32)
    for j in range(f):
        if j < n_sparse:
            nz = rng.random(n, dtype=np.float32) >= np.float32(0.95)
            col = np.zeros(n, np.float32)
            col[nz] = rng.integers(1, 64, int(nz.sum())).astype(np.float32)
        else:
            col = rng.integers(0, 64, n).astype(np.float32)
        X[:, j] = col
    trng = _rng(config_id, 50, 3, 7)
    y = np.zeros(n, np.float32)
    for _ in range(50):
        feats = trng.choice(f, 3, replace=False)
        thr = trng.integers(1, 63, 3)
        w = trng.standard_normal(8).astype(np.float32)
        idx = np.zeros(n, np.int64)
        for k in range(3):
            idx = idx * 2 + (X[:, feats[k]] > thr[k])
        y += w[idx]
    y += np.float32(0.1) * _rng(config_id, n, f, 200 + stream).standard_normal(n, dtype=np.float32)
    return X, y.astype(np.float32)


def auc(scores: np.ndarray, labels: np.ndarray) -> float:
    """Mann-Whitney rank-sum AUC with average ranks for ties (src/metric.cpp:95-137)."""
    s = np.asarray(scores, np.float64)
    order = np.argsort(s, kind="mergesort")
    ss = s[order]
    n = s.size
    # average 1-based rank per tie group
    bounds = np.flatnonzero(np.diff(ss)) + 1
    starts = np.concatenate([[0], bounds])
    ends = np.concatenate([bounds, [n]])
    ranks = np.empty(n, np.float64)
    avg = 0.5 * (starts + 1 + ends)
    ranks[order] = np.repeat(avg, ends - starts)
    pos = np.asarray(labels) >= 0.5
    n_pos = int(pos.sum())
    n_neg = n - n_pos
    if n_pos == 0 or n_neg == 0:
        return 1.0
    u = ranks[pos].sum() - n_pos * (n_pos + 1) / 2.0
    return float(u / (n_pos * n_neg))
