"""C ABI checks that need no GPU: the library loads, exports every function that
include/hjm_cva.h declares, its host-side utilities agree with the oracle, and invalid
arguments are rejected with status codes before any device work."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import cva_hjm_cuda_amd as hjm
from cva_hjm_cuda_amd import api

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for fn in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
            text = re.sub(r"//[^\n]*", " ", text)
            names |= set(re.findall(r"\b(hjm_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_library_exports_all_declared_symbols():
    lib = hjm.library()
    names = declared_functions()
    assert len(names) >= 12
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_status_strings():
    lib = hjm.library()
    for code in range(9):
        assert lib.hjm_status_string(code)
    assert b"grid" in lib.hjm_status_string(3)


def test_drift_matches_oracle(oracle):
    import json
    mk = json.load(open(os.path.join(ROOT, "tests", "golden", "legacy_market.json")))
    v = np.array(mk["volatilities"])
    assert np.allclose(hjm.hjm_drift(v), oracle.hjm_drift(v), rtol=0, atol=1e-18)
    assert np.array_equal(np.round(hjm.hjm_drift(v), 7), np.array(mk["drifts"]))


def test_cva_weights_match_oracle(oracle):
    for h, R in ((0.02, 0.4), (0.05, 0.0), (0.0, 0.4)):
        assert np.allclose(hjm.cva_weights(h, R), oracle.cva_weights(51, 0.5, h, R), rtol=1e-15,
                           atol=0)


def test_benchmark_market_matches_oracle(oracle):
    m = hjm.benchmark_market()
    assert np.allclose(m.spot_rates, oracle.nelson_siegel(), rtol=1e-15, atol=0)
    assert np.allclose(m.volatilities, oracle.poly_vols(), rtol=0, atol=1e-16)
    assert np.allclose(m.drifts, oracle.hjm_drift(oracle.poly_vols()), rtol=0, atol=1e-18)


def _args(paths=16, expiry=5.0, dt=0.01, normals=None):
    m = hjm.benchmark_market()
    mk = m._c()
    sw = api._Swap(1e6, 0.025, expiry, 0.5)
    cfg = api._config(paths, dt, hjm.FP64, api.RNG_PHILOX if normals is None else api.RNG_HOST,
                      1, normals, 0.02, 0.4, 0.975, 0, 0)
    res = api._Results()
    return m, mk, sw, cfg, res


@pytest.mark.parametrize("case,code", [
    ("null_market", 1), ("null_results", 1), ("paths0", 2), ("dt_not_dividing", 3),
    ("partial_block", 3), ("beyond_grid", 3), ("bad_recovery", 2), ("host_rng_no_normals", 1),
    ("bad_precision", 2)])
def test_argument_errors_before_device(case, code):
    lib = hjm.library()
    m, mk, sw, cfg, res = _args()
    mkp, resp = C.byref(mk), C.byref(res)
    if case == "null_market":
        mkp = None
    elif case == "null_results":
        resp = None
    elif case == "paths0":
        cfg.paths = 0
    elif case == "dt_not_dividing":
        cfg.dt = 0.03
    elif case == "partial_block":
        sw.expiry = 5.25
    elif case == "beyond_grid":
        sw.expiry = 26.0
    elif case == "bad_recovery":
        cfg.recovery = 1.5
    elif case == "host_rng_no_normals":
        cfg.rng = api.RNG_HOST
    elif case == "bad_precision":
        cfg.precision = 9
    assert lib.hjm_simulate_reference(mkp, C.byref(sw), C.byref(cfg), resp) == code


def test_nan_market_rejected():
    lib = hjm.library()
    m, mk, sw, cfg, res = _args()
    m.volatilities[5] = np.nan
    assert lib.hjm_simulate_reference(C.byref(mk), C.byref(sw), C.byref(cfg), C.byref(res)) == 2


def test_dropin_argument_errors():
    lib = hjm.library()
    m = hjm.benchmark_market()
    ee = np.zeros(51)
    p = api._ptr
    assert lib.hjm_calculate_exposure_multi_gpu_f64(None, 1.0, 0.02, 5.0, 0.5, p(m.accrual),
                                                    p(m.spot_rates), p(m.drifts),
                                                    p(m.volatilities), 0.0, 1, 100, 0.01) == 1
    assert lib.hjm_calculate_exposure_multi_gpu_f64(p(ee), 1.0, 0.02, 5.0, 0.5, p(m.accrual),
                                                    p(m.spot_rates), p(m.drifts),
                                                    p(m.volatilities), 0.0, 0, 100, 0.01) == 2


def test_no_cpu_fallback_without_device():
    """On a host without a GPU the product reports a device error; it never computes."""
    if hjm.device_count() > 0 if _has_device() else False:
        pytest.skip("a GPU is present")
    m = hjm.benchmark_market()
    with pytest.raises(hjm.HJMError) as ei:
        hjm.simulate_reference(m, hjm.InterestRateSwap(1e6, 0.025, 5.0), 64)
    assert ei.value.code == 4


def _has_device():
    try:
        return hjm.device_count() > 0
    except hjm.HJMError:
        return False


@pytest.mark.parametrize("shape", [(16, 499, 3), (15, 500, 3), (16, 500, 2), (16 * 500 * 3,)])
def test_host_normals_shape_rejected_before_the_library(shape):
    """Every entry point that takes caller normals checks their shape in Python: the C
    side copies exactly paths x steps x 3 values from the pointer."""
    from cva_hjm_cuda_amd.portfolio import synthetic_portfolio
    m = hjm.benchmark_market()
    swap = hjm.InterestRateSwap(1e6, 0.025, 5.0)
    z = np.zeros(shape)
    with pytest.raises(ValueError):
        hjm.simulate_reference(m, swap, 16, normals=z)
    with pytest.raises(ValueError):
        hjm.cva_greeks(m, swap, 16, normals=z)
    pf = synthetic_portfolio(4, 2, seed=1)
    with pytest.raises(ValueError):
        hjm.PortfolioPlan(m, pf, 16, horizon=5.0, normals=z)
    with pytest.raises(ValueError):
        hjm.portfolio_cva_greeks(m, pf, 16, horizon=5.0, normals=z)


_PF_CASES = [
    ("null_portfolio", 1), ("no_swaps", 2), ("no_sets", 2), ("period_zero", 2),
    ("period_past_horizon", 2), ("set_out_of_range", 2), ("nan_notional", 2),
    ("negative_hazard", 2), ("recovery_above_one", 2), ("horizon_not_whole", 3),
    ("horizon_past_30y", 3), ("dt_not_dividing", 3), ("paths0", 2), ("quantile_above_one", 2),
    ("bad_precision", 2), ("bad_path_kernel", 2), ("bad_adjoint_normals", 2),
    ("host_rng_no_normals", 1), ("matrix_with_host_normals", 7), ("bad_block_size", 2),
    ("bad_path_shape", 2), ("bad_value_shape", 2), ("bad_value_stage", 2),
    ("bad_stage_slots", 2)]


@pytest.mark.parametrize("case,code", _PF_CASES)
def test_portfolio_argument_errors_before_device(case, code):
    """hjm_portfolio_plan_create checks the book (periods within the horizon, sets in range,
    finite notionals and rates, non-negative hazards), the grid (a whole number of periods,
    at most 60, dt dividing dtau), the configuration and the tuning before any device work,
    and leaves the plan handle NULL."""
    lib = hjm.library()
    m = hjm.benchmark_market()
    mk = m._c()
    pf = hjm.synthetic_portfolio(8, 3, seed=1)
    periods, notional, sets, hazard = (pf.periods.copy(), pf.sign_notional.copy(),
                                       pf.set_of.copy(), pf.hazard.copy())
    pf = hjm.Portfolio(np.minimum(periods, 20), notional, pf.fixed_rate, sets, hazard)
    pfc = api._pf_c(pf, 10.0, 0.4)
    z = np.zeros((16, 1000, 3))
    cfg = api._config(16, 0.01, hjm.FP32_KAHAN, api.RNG_PHILOX, 1, None, 0.0, 0.4, 0.975, 0, 0)
    tun = hjm.Tuning()
    pfp = C.byref(pfc)
    if case == "null_portfolio":
        pfp = None
    elif case == "no_swaps":
        pfc.nswaps = 0
    elif case == "no_sets":
        pfc.nsets = 0
    elif case == "period_zero":
        pf.periods[0] = 0
    elif case == "period_past_horizon":
        pf.periods[0] = 21          # 10y of 0.5y periods is 20
    elif case == "set_out_of_range":
        pf.set_of[0] = pf.nsets
    elif case == "nan_notional":
        pf.sign_notional[0] = np.nan
    elif case == "negative_hazard":
        pf.hazard[0] = -0.01
    elif case == "recovery_above_one":
        pfc.recovery = 1.5
    elif case == "horizon_not_whole":
        pfc.horizon = 10.25
    elif case == "horizon_past_30y":
        pfc.horizon = 30.5          # 61 periods
    elif case == "dt_not_dividing":
        cfg.dt = 0.03
    elif case == "paths0":
        cfg.paths = 0
    elif case == "quantile_above_one":
        cfg.pfe_quantile = 1.5
    elif case == "bad_precision":
        cfg.precision = 9
    elif case == "bad_path_kernel":
        cfg.path_kernel = 9
    elif case == "bad_adjoint_normals":
        cfg.adjoint_normals = 5
    elif case == "host_rng_no_normals":
        cfg.rng = api.RNG_HOST
    elif case == "matrix_with_host_normals":
        cfg = api._config(16, 0.01, hjm.FP32_KAHAN, api.RNG_HOST, 1, z, 0.0, 0.4, 0.975, 0, 0,
                          path_kernel=hjm.KERNEL_MATRIX)
    elif case == "bad_block_size":
        tun.block_size = 100
    elif case == "bad_path_shape":
        tun.path_tiles = 3
    elif case == "bad_value_shape":
        tun.value_waves = 5
    elif case == "bad_value_stage":
        tun.value_stage = 9
    elif case == "bad_stage_slots":
        tun.pfe_stage_slots = 33
    tc = tun._c()
    h = C.c_void_p(1)
    assert lib.hjm_portfolio_plan_create(C.byref(h), C.byref(mk), pfp, C.byref(cfg),
                                         C.byref(tc)) == code
    assert not h.value
