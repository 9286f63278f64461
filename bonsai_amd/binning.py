"""Host bin mapping (H1): cut fitting and the bin transform.

Follows the reference's BinMapper (src/bin_mapper.cpp:57-291): NaN-free working set,
sort + run-length encode, then one cut per distinct value when they fit the budget,
the greedy heavy-value rule when some value outweighs a mean bin, else the quantile
stride; every path appends the FLT_MAX closer and the +inf missing sentinel, so
n_bins = len(cuts) <= max_bin.  transform = lower_bound, NaN -> last bin.

Row sampling above n_samples is this package's own (numpy PCG64 keyed on the seed):
the reference draws its sample with std::ranges::sample + std::mt19937, whose index
selection is standard-library specific (src/bin_mappers.cpp:31-45).
"""
from __future__ import annotations

import numpy as np

FLT_MAX = np.float32(np.finfo(np.float32).max)


def _midpoint(a: np.float32, b: np.float32) -> np.float32:
    # std::midpoint for floats: (a + b) / 2 rounded once when no overflow is possible.
    a = np.float32(a)
    b = np.float32(b)
    hi = np.float32(FLT_MAX / np.float32(2))
    if abs(a) <= hi and abs(b) <= hi:
        return np.float32((a + b) / np.float32(2))
    return np.float32(a / np.float32(2) + b / np.float32(2))


def _greedy(vals, counts, n_samples, cut_budget, mean_bin):
    cuts: list[np.float32] = []
    n_groups = cut_budget + 1
    is_big = counts.astype(np.float64) >= mean_bin
    n_big = int(is_big.sum())
    big_sum = int(counts[is_big].sum())
    rest_sum = n_samples - big_sum
    rest_groups = n_groups - n_big
    bin_size = rest_sum / rest_groups if rest_groups != 0 else mean_bin
    in_bin = 0
    i = 0
    while i + 1 < len(vals) and len(cuts) < cut_budget:
        if not is_big[i]:
            rest_sum -= int(counts[i])
        in_bin += int(counts[i])
        if is_big[i] or in_bin >= bin_size or (is_big[i + 1] and in_bin >= bin_size / 2.0):
            mid = _midpoint(vals[i], vals[i + 1])
            if not cuts or cuts[-1] < mid:
                cuts.append(mid)
            in_bin = 0
            if not is_big[i] and rest_groups > 1:
                rest_groups -= 1
                bin_size = rest_sum / rest_groups
        i += 1
    return np.asarray(cuts, np.float32)


def create_cuts(sample: np.ndarray, max_bin: int) -> np.ndarray:
    """Cuts from a NaN-free working set (BinMapper::from_sample)."""
    if max_bin <= 2:
        raise ValueError("max_bin must exceed 2")
    cut_budget = max_bin - 2
    s = np.sort(np.asarray(sample, np.float32), kind="stable")
    n = s.size
    if n:
        vals, counts = np.unique(s, return_counts=True)
    else:
        vals, counts = np.zeros(0, np.float32), np.zeros(0, np.int64)
    mean_bin = n / (cut_budget + 1)
    if vals.size <= cut_budget:
        cuts = vals.astype(np.float32)
    elif counts.max() >= mean_bin:
        cuts = _greedy(vals, counts, n, cut_budget, mean_bin)
    else:
        step = max((n + cut_budget) // (cut_budget + 1), 1)
        picked = s[np.arange(step, n, step)]
        cuts = np.unique(picked).astype(np.float32)
    out = list(cuts)
    if not out or out[-1] < FLT_MAX:
        out.append(FLT_MAX)
    out.append(np.float32(np.inf))
    return np.asarray(out, np.float32)


def sample_rows(n: int, n_samples: int = 200_000, seed: int = 0) -> np.ndarray | None:
    """Ascending sample rows the cuts are fitted on (None: every row)."""
    if n > n_samples:
        return np.sort(np.random.default_rng(seed).choice(n, n_samples, replace=False))
    return None


def cuts_of_sample(S: np.ndarray, max_bin: int) -> list[np.ndarray]:
    cuts = []
    for j in range(S.shape[1]):
        col = S[:, j]
        cuts.append(create_cuts(col[~np.isnan(col)], max_bin))
    return cuts


def fit_cuts(X: np.ndarray, max_bin: int = 255, n_samples: int = 200_000, seed: int = 0):
    """Per-feature cuts for a row-major float matrix."""
    rows = sample_rows(X.shape[0], n_samples, seed)
    return cuts_of_sample(X if rows is None else X[rows], max_bin)


def fit_cuts_device(X, max_bin: int = 255, n_samples: int = 200_000, seed: int = 0):
    """fit_cuts for a device-resident matrix (bonsai_amd.DeviceMatrix): only the sample
    rows are gathered on the device and copied to the host (cuda_gather_rows,
    src/cuda/histogram_engine.cu:473-493).  Same cuts as fit_cuts on the same matrix."""
    from bonsai_amd import gather_rows
    rows = sample_rows(X.n_rows, n_samples, seed)
    return cuts_of_sample(gather_rows(X, rows), max_bin)


def transform(X: np.ndarray, cuts: list[np.ndarray]) -> np.ndarray:
    """Row-major u8 bins: lower_bound over the cuts, NaN -> last bin."""
    n, f = X.shape
    out = np.empty((n, f), np.uint8)
    for j in range(f):
        c = cuts[j]
        if c.size > 256:
            raise ValueError("more than 256 bins: u8 planes only")
        col = X[:, j]
        b = np.searchsorted(c, col, side="left")
        b[np.isnan(col)] = c.size - 1
        out[:, j] = b.astype(np.uint8)
    return out


def n_bins_of(cuts: list[np.ndarray]) -> np.ndarray:
    return np.asarray([c.size for c in cuts], np.uint16)
