"""Pins of the calibration oracle (oracle/calibration.py) by constructions with a known
answer, no GPU.

* Differences built as Z sigma with centred, scaled-orthonormal Z have annualised sample
  covariance exactly sigma^T sigma: the PCA's top three eigenvalues are those of the 3x3
  Gram matrix sigma sigma^T, the rest vanish, and the top three loadings reproduce the
  covariance.
* The HJM drift depends on the loadings only through the covariance, so the drift of the
  PCA loadings equals the drift of the generating loadings (benchmark market), although
  the loadings themselves differ (a rotation); dropping a factor breaks it.
* The benchmark loadings are polynomials of degree <= 3 in the tenor, so a cubic fit of
  any PCA loading is exact; a quadratic fit is not.
* Sign convention and rolling windows.
"""
import os
import sys

import numpy as np
import pytest

import cva_hjm_cuda_amd as hjm

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                "oracle"))
import calibration as cal  # noqa: E402


@pytest.fixture(scope="module")
def sigma():
    return hjm.benchmark_market().volatilities.reshape(3, 51)


def test_rank3_history_has_the_generating_covariance(sigma):
    curves = hjm.synthetic_history(sigma, 600, seed=3)
    lam, raw, C = cal.pca_window(curves)
    G = sigma.T @ sigma
    assert np.max(np.abs(C - G)) < 1e-12 * np.max(np.abs(G))
    gram = np.sort(np.linalg.eigvalsh(sigma @ sigma.T))[::-1]
    np.testing.assert_allclose(lam[:3], gram, rtol=1e-10)
    assert np.max(np.abs(lam[3:])) < 1e-12 * lam[0]
    assert np.max(np.abs(raw.T @ raw - G)) < 1e-12 * np.max(np.abs(G))
    # loadings are orthogonal (PCA), the generating ones are not
    off = raw @ raw.T - np.diag(np.diag(raw @ raw.T))
    assert np.max(np.abs(off)) < 1e-12 * lam[0]
    assert abs((sigma @ sigma.T)[0, 1]) > 1e-6 * lam[0]


def test_drift_of_pca_loadings_equals_generating_drift(sigma, oracle):
    curves = hjm.synthetic_history(sigma, 500, seed=5)
    _, raw, _ = cal.pca_window(curves)
    m_gen = oracle.hjm_drift(sigma.reshape(-1))
    m_pca = oracle.hjm_drift(raw.reshape(-1))
    assert np.max(np.abs(m_pca - m_gen)) < 1e-10 * np.max(np.abs(m_gen))
    assert np.max(np.abs(raw - sigma)) > 1e-4       # a rotation, not the same loadings
    two = raw.copy()
    two[2] = 0.0
    assert np.max(np.abs(oracle.hjm_drift(two.reshape(-1)) - m_gen)) > 1e-3 * np.max(np.abs(m_gen))


def test_cubic_fit_is_exact_quadratic_is_not(sigma):
    curves = hjm.synthetic_history(sigma, 400, seed=7)
    _, raw, _ = cal.pca_window(curves)
    np.testing.assert_allclose(cal.fit(raw, (3, 3, 3)), raw, rtol=0, atol=1e-12)
    assert np.max(np.abs(cal.fit(raw, (2, 2, 2)) - raw)) > 1e-6
    np.testing.assert_allclose(cal.fit(raw, (-1, 3, -1))[[0, 2]], raw[[0, 2]], rtol=0, atol=0)


def test_sign_convention_and_windows(sigma, oracle):
    curves = hjm.synthetic_history(sigma, 300, seed=11, noise=1e-5)
    ev, vols, drifts, raws = cal.calibrate(curves, 120, 30, 6, 252.0, None, 0.5,
                                           lambda v, h: oracle.hjm_drift(v, h))
    for w in range(6):
        lam, raw, _ = cal.pca_window(curves[30 * w:30 * w + 120])
        np.testing.assert_array_equal(ev[w], lam)
        np.testing.assert_array_equal(raws[w], raw.reshape(-1))
        for k in range(3):
            assert raw[k].sum() > 0
    assert np.all(np.diff(ev, axis=1) <= 0)
    assert np.array_equal(cal.fix_sign(-np.arange(51.0)), np.arange(51.0))
    z = np.zeros(51)
    z[3], z[4] = -2.0, 2.0
    assert np.array_equal(cal.fix_sign(z), -z)
