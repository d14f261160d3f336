"""Pins the oracle's EE / PFE / CVA / SE statistics.

* constant EE: CVA = (1-R) E (1 - e^{-h T}) exactly (telescoping default probabilities);
* PFE: numpy.quantile(method="inverted_cdf") on the same samples;
* EE and SE: numpy mean / sample standard deviation of the per-path CVA.
"""
import math

import numpy as np
import pytest


def test_cva_weights_constant_exposure(oracle):
    h, R, dtd, n = 0.02, 0.4, 0.5, 51
    w = oracle.cva_weights(n, dtd, h, R)
    T = (n - 1) * dtd
    assert w.sum() == pytest.approx((1 - R) * (1 - math.exp(-h * T)), rel=1e-14)
    assert np.all(w > 0)
    # trapezoid: end weights are half an interval's default probability
    pd1 = math.exp(0.0) - math.exp(-h * dtd)
    assert w[0] == pytest.approx(0.5 * (1 - R) * pd1, rel=1e-14)


def test_cva_zero_hazard(oracle):
    assert np.all(oracle.cva_weights(51, 0.5, 0.0, 0.4) == 0.0)


@pytest.mark.parametrize("n", [1, 2, 39, 40, 41, 1000, 1001, 4096])
def test_pfe_matches_numpy_inverted_cdf(oracle, n):
    rng = np.random.default_rng(n)
    ex = np.maximum(rng.normal(0.1, 1.0, size=(n, 51)), 0.0)
    w = oracle.cva_weights()
    ee, pfe, cva, se = oracle.stats(ex, w, 0.975)
    assert np.array_equal(pfe, np.quantile(ex, 0.975, axis=0, method="inverted_cdf"))
    assert np.allclose(ee, ex.mean(axis=0), rtol=1e-13, atol=0)
    per_path = ex @ w
    assert cva == pytest.approx(per_path.mean(), rel=1e-12)
    if n > 1:
        assert se == pytest.approx(per_path.std(ddof=1) / math.sqrt(n), rel=1e-9)


def test_quantile_rank(oracle):
    assert oracle.quantile_rank(1000, 0.975) == 975
    assert oracle.quantile_rank(131072, 0.975) == 127796
    assert oracle.quantile_rank(1, 0.975) == 1
    assert oracle.quantile_rank(10, 1.0) == 10
    assert oracle.quantile_rank(10, 0.0) == 1
