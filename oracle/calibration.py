"""CPU oracle of the PCA calibration upstream of the path (test infrastructure only: loaded
by tests/ and bench.py's CPU baseline leg, never by the library).

The reference holds no calibration code; its roadmap names the step (doc/TODO:16, "Add the
volatility and drift for HJM model calibration"). This restates the classic HJM PCA
calibration the library implements, in numpy:
  d_t = f_{t+1} - f_t over a window of curves; C = obs_per_year * cov(d) (ddof 1);
  C = sum_i lambda_i v_i v_i^T (numpy.linalg.eigh), lambda descending;
  sign of v_k: sum of components > 0, else the first component above 1e-9 positive;
  sigma_k = sqrt(max(lambda_k, 0)) v_k; optional least-squares polynomial of degree d_k in
  x = j / 50 (numpy.polyfit); drift m = cumulative Simpson (the C oracle's orc_hjm_drift,
  passed in as drift_fn).
Pinned in tests/test_oracle_calibration.py by constructions with a known answer.
"""
import numpy as np

TENORS = 51


def fix_sign(v):
    s = v.sum()
    if abs(s) > 1e-9:
        return v if s > 0 else -v
    for x in v:
        if abs(x) > 1e-9:
            return v if x > 0 else -v
    return v


def pca_window(curves, obs_per_year=252.0, nfactors=3):
    d = np.diff(np.asarray(curves, np.float64), axis=0)
    C = np.cov(d, rowvar=False, ddof=1) * obs_per_year
    lam, vec = np.linalg.eigh(C)
    order = np.argsort(lam)[::-1]
    lam, vec = lam[order], vec[:, order]
    raw = np.stack([np.sqrt(max(lam[k], 0.0)) * fix_sign(vec[:, k]) for k in range(nfactors)])
    return lam, raw, C


def fit(raw, degrees):
    x = np.arange(TENORS) / (TENORS - 1.0)
    out = raw.copy()
    for k, d in enumerate(degrees):
        if d is not None and d >= 0:
            out[k] = np.polyval(np.polyfit(x, raw[k], d), x)
    return out


def calibrate(curves, window, step, nwindows, obs_per_year, fit_degrees, dtau, drift_fn):
    """Returns eigenvalues [W][51], volatilities [W][153], drifts [W][51], raw [W][153]."""
    ev, vols, drifts, raws = [], [], [], []
    for w in range(nwindows):
        lam, raw, _ = pca_window(curves[w * step:w * step + window], obs_per_year)
        v = fit(raw, fit_degrees) if fit_degrees is not None else raw
        ev.append(lam)
        raws.append(raw.reshape(-1))
        vols.append(v.reshape(-1))
        drifts.append(drift_fn(v.reshape(-1), dtau))
    return np.array(ev), np.array(vols), np.array(drifts), np.array(raws)
