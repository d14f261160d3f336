"""Pins the oracle's FP64 pathwise adjoint (CVA Greeks, reference semantics) against
central finite differences of the oracle's own CVA with common random numbers, for all
51 + 153 + 51 parameters, and the drift Jacobian-transpose product against finite
differences of the Simpson drift."""
import numpy as np
import pytest

T = 51


@pytest.fixture(scope="module")
def setup(oracle):
    spot = oracle.nelson_siegel()
    vols = oracle.poly_vols()
    drift = oracle.hjm_drift(vols)
    acc = np.full(T, 0.5)
    z = oracle.normals(11, 64, 300)
    w = oracle.cva_weights(T, 0.5, 0.02, 0.4)
    args = dict(accrual=acc, notional=1e6, K=0.025, expiry=3.0, dtau=0.5, dt=0.01, normals=z,
                w=w)
    return spot, drift, vols, args


def _fd(oracle, spot, drift, vols, args, which, j, h):
    def cva(s, d, v):
        return oracle.reference_cva(s, d, v, **args)
    arrs = {"spot": spot, "drift": drift, "vols": vols}
    up = {k: a.copy() for k, a in arrs.items()}
    dn = {k: a.copy() for k, a in arrs.items()}
    up[which][j] += h
    dn[which][j] -= h
    return (cva(up["spot"], up["drift"], up["vols"]) - cva(dn["spot"], dn["drift"],
                                                           dn["vols"])) / (2 * h)


def test_adjoint_matches_central_fd(oracle, setup):
    spot, drift, vols, args = setup
    cva, ds, dv, dm = oracle.reference_cva_greeks(spot, drift, vols, **args)
    assert cva == pytest.approx(oracle.reference_cva(spot, drift, vols, **args), rel=1e-12)
    for which, g, h in (("spot", ds, 1e-6), ("drift", dm, 1e-7), ("vols", dv, 1e-7)):
        fd = np.array([_fd(oracle, spot, drift, vols, args, which, j, h) for j in range(len(g))])
        scale = np.max(np.abs(fd))
        assert scale > 0
        assert np.max(np.abs(g - fd)) < 2e-5 * scale, (which, np.max(np.abs(g - fd)) / scale)


def test_adjoint_detects_planted_sign_error(oracle, setup):
    """Negative control: a sign flip in one Greek is far outside the FD tolerance."""
    spot, drift, vols, args = setup
    _, ds, _, _ = oracle.reference_cva_greeks(spot, drift, vols, **args)
    j = int(np.argmax(np.abs(ds)))
    fd = _fd(oracle, spot, drift, vols, args, "spot", j, 1e-6)
    assert abs(-ds[j] - fd) > 0.5 * abs(fd)


def test_drift_vjp_matches_fd(oracle):
    vols = oracle.poly_vols()
    rng = np.random.default_rng(3)
    g = rng.normal(size=T)
    got = oracle.drift_vjp(vols, g)
    h = 1e-7
    for j in rng.choice(3 * T, 20, replace=False):
        up, dn = vols.copy(), vols.copy()
        up[j] += h
        dn[j] -= h
        fd = (oracle.hjm_drift(up) @ g - oracle.hjm_drift(dn) @ g) / (2 * h)
        assert got[j] == pytest.approx(fd, rel=1e-6, abs=1e-12)
