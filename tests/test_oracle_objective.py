"""The oracle's objective functions (src/objective.cpp:44-130) against closed forms and
the reference's own eval cases (tests/unit/test_objective.cpp:82-250).

Each check is one a plausible slip fails: a flipped sign or a swapped p / 1 - p in the
logistic gradient or the init score, a dropped softplus term in the eval, a mean taken
over the wrong count.
"""
import math

import numpy as np
import pytest

MSE, LOGLOSS = 0, 1


def test_logloss_grad_matches_the_double_closed_form(orc):
    """g = sigmoid(s) - y, h = sigmoid(s)(1 - sigmoid(s)) within float rounding of the
    double-precision closed form; h in [0, 0.25]; s = 0 gives exactly 0.5 - y and 0.25."""
    rng = np.random.default_rng(0)
    s = np.concatenate([rng.normal(0, 3, 4000), [0.0, 0.0, 30.0, -30.0, 88.0, -88.0]]).astype(np.float32)
    y = (rng.random(s.size) < 0.5).astype(np.float32)
    g, h = orc.logloss_grad(s, y)
    p = 1.0 / (1.0 + np.exp(-s.astype(np.float64)))
    assert np.max(np.abs(g - (p - y))) <= 2e-7
    assert np.max(np.abs(h - p * (1.0 - p))) <= 2e-7
    assert np.all((h >= 0) & (h <= 0.25))
    z = s == 0
    assert np.array_equal(g[z], (np.float32(0.5) - y[z]))
    assert np.all(h[z] == np.float32(0.25))


def test_mse_grad_is_residual_and_unit_hessian(orc):
    rng = np.random.default_rng(1)
    s = rng.normal(0, 2, 1000).astype(np.float32)
    y = rng.normal(0, 2, 1000).astype(np.float32)
    g, h = orc.mse_grad(s, y)
    assert np.array_equal(g, s - y)
    assert np.all(h == 1.0)


def test_init_score_mse_is_the_float_mean(orc):
    """MSEObjective::init_score (src/objective.cpp:77-84): sum / n in float."""
    assert orc.init_score(MSE, np.array([1, 2, 3, 4], np.float32)) == 2.5
    assert orc.init_score(MSE, np.array([-3, -3, -3], np.float32)) == -3.0
    rng = np.random.default_rng(2)
    y = rng.normal(1.5, 1.0, 5000).astype(np.float32)
    assert orc.init_score(MSE, y) == pytest.approx(math.fsum(y.astype(np.float64)) / y.size, rel=1e-5)


def test_init_score_logloss_is_the_log_odds(orc):
    """LogLossObjective::init_score (src/objective.cpp:122-130): log(p / (1 - p)) of the
    label mean; p = 3/4 gives log 3, p = 1/2 gives 0, and swapping p and 1 - p flips the sign."""
    assert orc.init_score(LOGLOSS, np.array([0, 1, 1, 1], np.float32)) == pytest.approx(math.log(3.0), abs=1e-6)
    assert orc.init_score(LOGLOSS, np.array([1, 0, 0, 0], np.float32)) == pytest.approx(-math.log(3.0), abs=1e-6)
    assert orc.init_score(LOGLOSS, np.array([0, 1] * 50, np.float32)) == 0.0
    rng = np.random.default_rng(3)
    y = (rng.random(20000) < 0.3).astype(np.float32)
    p = y.astype(np.float64).mean()
    assert orc.init_score(LOGLOSS, y) == pytest.approx(math.log(p / (1 - p)), abs=1e-5)


def test_logloss_eval_reference_cases(orc):
    """The reference's three eval cases (tests/unit/test_objective.cpp:211-254)."""
    # all scores 0: log 2 whatever the labels
    assert orc.logloss_eval(np.zeros(4, np.float32), np.array([0, 1, 0, 1], np.float32)) == pytest.approx(
        math.log(2.0), abs=1e-6)
    # against the double stable-softplus cross entropy
    s = np.array([0.0, 1.0, -1.0, 2.0, -2.0], np.float32)
    y = np.array([1.0, 0.0, 1.0, 1.0, 0.0], np.float32)
    ref = sum(max(0.0, float(x)) + math.log1p(math.exp(-abs(float(x)))) - float(t) * float(x) for x, t in zip(s, y))
    assert orc.logloss_eval(s, y) == pytest.approx(ref / s.size, abs=1e-6)
    # extreme, correctly classified scores: finite and near zero
    loss = orc.logloss_eval(np.array([100, -100, 1000, -1000], np.float32), np.array([1, 0, 1, 0], np.float32))
    assert math.isfinite(loss) and abs(loss) <= 1e-3


def test_logloss_eval_is_the_mean_cross_entropy(orc):
    """Random scores against -y log p - (1 - y) log(1 - p) in double."""
    rng = np.random.default_rng(4)
    s = rng.normal(0, 2, 3000).astype(np.float32)
    y = (rng.random(s.size) < 0.5).astype(np.float32)
    p = 1.0 / (1.0 + np.exp(-s.astype(np.float64)))
    ce = -(y * np.log(p) + (1 - y) * np.log1p(-p)).mean()
    assert orc.logloss_eval(s, y) == pytest.approx(ce, rel=1e-5)


def test_poisson_grad_closed_form_and_clamp(orc):
    """PoissonObjective::compute (src/objective.cpp:253-278): g = exp(s) - y, h = exp(s)
    with the raw score clamped to +-30; s = 0 gives exactly 1 - y and 1."""
    rng = np.random.default_rng(5)
    s = np.concatenate([rng.normal(0, 2, 2000), [0.0, 31.0, 100.0, -31.0, -100.0]]).astype(np.float32)
    y = rng.poisson(2.0, s.size).astype(np.float32)
    g, h = orc.poisson_grad(s, y)
    mu = np.exp(np.clip(s.astype(np.float64), -30.0, 30.0))
    assert np.all(np.abs(h - mu) <= 2e-7 * mu)
    assert np.all(np.abs(g - (mu - y)) <= 2e-7 * mu + 1e-6 * np.abs(y))
    z = s == 0
    assert np.array_equal(g[z], np.float32(1.0) - y[z]) and np.all(h[z] == 1.0)
    assert h[s.size - 4] == h[s.size - 3] == np.float32(np.exp(np.float32(30.0)))   # clamped above
    assert h[s.size - 2] == h[s.size - 1] == np.float32(np.exp(np.float32(-30.0)))  # and below


def test_init_score_poisson_is_the_log_mean(orc):
    """PoissonObjective::init_score (src/objective.cpp:293-308): log of the double mean,
    floored at 1e-9 so all-zero labels stay finite."""
    assert orc.init_score(2, np.array([1, 2, 3, 6], np.float32)) == pytest.approx(math.log(3.0), abs=1e-6)
    assert orc.init_score(2, np.zeros(10, np.float32)) == pytest.approx(math.log(1e-9), rel=1e-6)
    rng = np.random.default_rng(6)
    y = rng.poisson(0.7, 10000).astype(np.float32)
    assert orc.init_score(2, y) == pytest.approx(math.log(y.astype(np.float64).mean()), abs=1e-6)
