"""Random material parameters (Weibull thresholds of the damage law) drawn as
the reference draws them, against the golden vector of
tests/golden/weibull_default_engine.txt (std::default_random_engine +
std::weibull_distribution, written by tests/golden/make_weibull_golden.py)."""
import math
import os

import numpy as np
import pytest

from akantu_amd import rng


def _golden(golden_dir):
    vals, seed = {}, None
    for line in open(os.path.join(golden_dir, "weibull_default_engine.txt")):
        if line.startswith("seed"):
            seed = int(line.split()[1])
            vals[seed] = []
        else:
            vals[seed].append(float(line))
    return vals


def test_weibull_thresholds_equal_golden(golden_dir):
    for seed, ref in _golden(golden_dir).items():
        got = rng.weibull_thresholds(len(ref), 1e6, 5.0, seed=seed)
        assert np.array_equal(got, np.array(ref)), seed


def test_weibull_moments():
    """Mean scale * Gamma(1 + 1/m), shape parameter order [scale, shape]."""
    x = rng.weibull_thresholds(200_000, 1e6, 5.0, seed=7)
    assert abs(x.mean() / (1e6 * math.gamma(1.2)) - 1) < 5e-3
    assert (x > 0).all()


def test_base_and_uniform():
    a = rng.random_values(10, "uniform", 2.0, 0.5, 1.5, seed=3)
    assert ((a >= 2.5) & (a < 3.5)).all()
    w0 = rng.weibull_thresholds(5, 1.0, 2.0, seed=11)
    w1 = rng.weibull_thresholds(5, 1.0, 2.0, seed=11, base=10.0)
    assert np.array_equal(w1, 10.0 + w0)
    with pytest.raises(ValueError):
        rng.random_values(3, "gamma")
