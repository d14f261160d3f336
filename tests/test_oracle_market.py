"""Pins the market-data precompute of the oracle.

* The cumulative Simpson HJM drift of the reference's volatility table reproduces the
  reference's drift table (simulation_gpu_testing.cpp:66-71 from :47-63) to its 7 printed
  decimals, and a plain trapezoid does not (so the check discriminates).
* Simpson is exact on cubics at every node; the Nelson-Siegel curve hits its limits.
"""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def legacy():
    with open(os.path.join(HERE, "golden", "legacy_market.json")) as f:
        return json.load(f)


def test_drift_table_reproduced(oracle):
    mk = legacy()
    vols = np.array(mk["volatilities"])
    drift = oracle.hjm_drift(vols)
    table = np.array(mk["drifts"])
    assert np.max(np.abs(drift - table)) <= 5.1e-8
    assert np.array_equal(np.round(drift, 7), table)


def test_drift_trapezoid_would_fail(oracle):
    """Negative control: the same construction with the trapezoid rule misses the table."""
    mk = legacy()
    vols = np.array(mk["volatilities"]).reshape(3, 51)
    h = 0.5
    m = np.zeros(51)
    for k in range(3):
        integ = np.concatenate([[0.0], np.cumsum(0.5 * h * (vols[k, 1:] + vols[k, :-1]))])
        m += vols[k] * integ
    assert not np.array_equal(np.round(m, 7), np.array(mk["drifts"]))


def test_cumulative_simpson_exact_on_cubics(oracle):
    x = np.arange(51) * 0.5
    y = 0.3 - 0.2 * x + 0.05 * x**2 - 0.001 * x**3
    exact = 0.3 * x - 0.1 * x**2 + 0.05 / 3 * x**3 - 0.00025 * x**4
    got = oracle.cumulative_simpson(y, 0.5)
    # j = 1 uses a quadratic, exact up to degree 2 only
    assert np.max(np.abs(got[2:] - exact[2:])) < 1e-12
    yq = 0.3 - 0.2 * x + 0.05 * x**2
    exq = 0.3 * x - 0.1 * x**2 + 0.05 / 3 * x**3
    assert abs(oracle.cumulative_simpson(yq, 0.5)[1] - exq[1]) < 1e-14


def test_poly_vol_drift_is_exact_integral(oracle):
    """The benchmark vols are polynomials of degree <= 3, so the drift is exact (j != 1)."""
    t = np.arange(51) * 0.5
    s = [np.full(51, 0.0064), -0.0036 + 0.0006 * t - 2e-5 * t**2,
         -0.0047 + 0.0018 * t - 1.5e-4 * t**2 + 3e-6 * t**3]
    I = [0.0064 * t, -0.0036 * t + 0.0003 * t**2 - 2e-5 / 3 * t**3,
         -0.0047 * t + 0.0009 * t**2 - 5e-5 * t**3 + 7.5e-7 * t**4]
    exact = sum(a * b for a, b in zip(s, I))
    vols = oracle.poly_vols()
    assert np.allclose(vols.reshape(3, 51), np.array(s), rtol=0, atol=1e-16)
    d = oracle.hjm_drift(vols)
    assert np.max(np.abs(d[2:] - exact[2:])) < 1e-15
    assert abs(d[1] - exact[1]) < 1e-9


def test_nelson_siegel_curve(oracle):
    f = oracle.nelson_siegel()
    assert abs(f[0] - 0.01) < 1e-15
    tau = 25.0
    e = np.exp(-tau / 2)
    assert abs(f[50] - (0.03 - 0.02 * e + 0.01 * tau / 2 * e)) < 1e-15
    # hump: rises from 1% past the 3% asymptote, then decays back towards it
    assert f.max() > 0.03 and abs(f[50] - 0.03) < 1e-6
    assert np.all(np.diff(f[:8]) > 0)
