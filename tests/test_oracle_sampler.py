"""The oracle's Bernoulli row sampler (BernoulliSampler::sample, src/sampler.cpp:38-55,
with the product's counter-based draw instead of std::mt19937).

Pinned by what the distribution fixes: p >= 1 keeps every row (the reference's own
branch), the kept count is Binomial(n, p) (checked within 6 standard deviations), draws
of different iterations and seeds are independent (the overlap of two draws is
Binomial(n, p^2)), rows come out ascending and unique, and a threshold of zero keeps
nothing.  The row subset itself is the input the growers take, so tree parity on a
sampled root is checked against the oracle's growers on the same rows.
"""
import numpy as np
import pytest


def within(k, n, p, sig=6.0):
    mu, sd = n * p, np.sqrt(n * p * (1.0 - p))
    return abs(k - mu) <= sig * sd


def test_keep_all_at_or_above_one(orc):
    for p in (1.0, 1.5):
        assert np.array_equal(orc.bernoulli_rows(1000, p, 3, 0), np.arange(1000, dtype=np.uint32))


@pytest.mark.parametrize("p", [0.05, 0.3, 0.5, 0.9])
def test_kept_count_is_binomial(orc, p):
    n = 200_000
    rows = orc.bernoulli_rows(n, p, 11, 4)
    assert within(rows.size, n, p)
    assert np.all(np.diff(rows.astype(np.int64)) > 0)
    assert rows.size == 0 or rows[-1] < n


def test_iterations_and_seeds_are_independent(orc):
    n, p = 200_000, 0.4
    a = orc.bernoulli_rows(n, p, 11, 0)
    b = orc.bernoulli_rows(n, p, 11, 1)
    c = orc.bernoulli_rows(n, p, 12, 0)
    for x, y in ((a, b), (a, c), (b, c)):
        assert within(np.intersect1d(x, y).size, n, p * p)
    # the same key draws the same rows
    assert np.array_equal(a, orc.bernoulli_rows(n, p, 11, 0))


def test_spread_over_the_rows(orc):
    # kept rows are uniform over the index range: each tenth holds about a tenth
    n, p = 200_000, 0.25
    rows = orc.bernoulli_rows(n, p, 5, 2)
    counts = np.bincount(rows // (n // 10), minlength=10)
    for k in counts:
        assert within(k, n // 10, p)


def test_tiny_probability_keeps_almost_nothing(orc):
    assert orc.bernoulli_rows(100_000, 1e-12, 1, 0).size == 0
