"""Synthetic histories of forward curves for the calibration (tests, benchmarks, demos)."""
import numpy as np

TENORS = 51


def synthetic_history(volatilities, n, obs_per_year=252.0, base=None, seed=0, noise=0.0):
    """n daily forward curves [n][51] whose differences have annualised sample covariance
    exactly sigma^T sigma (sigma = volatilities as [3][51]) when noise = 0: differences
    Z sigma with Z centred and Z^T Z = (n - 2) / obs_per_year * I over the n - 1 days;
    ``noise`` adds i.i.d. N(0, noise^2) to every difference. ``base``: the first curve
    (default flat 3%)."""
    rng = np.random.default_rng(seed)
    m = n - 1
    Z = rng.standard_normal((m, 3))
    Z -= Z.mean(axis=0)
    q, _ = np.linalg.qr(Z)  # orthonormal columns in the span of centred ones: centred
    Z = q * np.sqrt((m - 1) / obs_per_year)
    d = Z @ np.asarray(volatilities, np.float64).reshape(3, TENORS)
    if noise:
        d = d + noise * rng.standard_normal(d.shape)
    f0 = np.full(TENORS, 0.03) if base is None else np.asarray(base, np.float64)
    return np.vstack([f0, f0 + np.cumsum(d, axis=0)])
