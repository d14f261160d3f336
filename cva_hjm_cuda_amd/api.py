"""Python host layer over libhjm_cva.so (the C ABI declared in include/hjm_cva.h).

Mirrors the reference's operator interface for the HJM exposure path:
``calculate_exposure_multi_gpu`` has the argument meaning of
``calculateExposureMultiGPU`` (simulation_gpu.h:10,12) and ``InterestRateSwap`` the
fields of product.h:4-17. Everything runs on the GPU through the HIP library; there is
no CPU fallback -- loading fails loudly if the library is missing.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "lib", "libhjm_cva.so")

TENORS = 51
FACTORS = 3
STAT_SLOTS = TENORS + 2

FP64 = 0
FP32_KAHAN = 1
RNG_PHILOX = 0
RNG_HOST = 1

STATUS = {
    0: "ok", 1: "null pointer", 2: "invalid argument", 3: "inconsistent time grid",
    4: "device error", 5: "allocation failed", 6: "launch failed", 7: "unsupported",
    8: "size beyond the call's fixed limit",
}
NOALLOC_MAX_PATHS = 32768   # HJM_NOALLOC_MAX_PATHS


class HJMError(RuntimeError):
    def __init__(self, code: int, what: str):
        super().__init__("%s failed: %s (status %d)" % (what, STATUS.get(code, "?"), code))
        self.code = code


_dp = C.POINTER(C.c_double)


class _Market(C.Structure):
    _fields_ = [("spot_rates", _dp), ("drifts", _dp), ("volatilities", _dp), ("accrual", _dp)]


class _Swap(C.Structure):
    _fields_ = [("notional", C.c_double), ("fixed_rate", C.c_double), ("expiry", C.c_double),
                ("dtau", C.c_double)]


class _Config(C.Structure):
    _fields_ = [("paths", C.c_int64), ("path_offset", C.c_int64), ("dt", C.c_double),
                ("precision", C.c_int32), ("rng", C.c_int32), ("seed", C.c_uint64),
                ("host_normals", _dp), ("hazard_rate", C.c_double), ("recovery", C.c_double),
                ("pfe_quantile", C.c_double), ("device", C.c_int32), ("path_kernel", C.c_int32),
                ("adjoint_normals", C.c_int32)]


class _Greeks(C.Structure):
    _fields_ = [("cva", _dp), ("cva_stderr", _dp), ("d_spot", _dp), ("d_vol", _dp),
                ("d_drift", _dp), ("d_vol_total", _dp), ("timings_ms", _dp),
                ("stored_normals_bytes", C.POINTER(C.c_int64))]


class _Portfolio(C.Structure):
    _fields_ = [("nswaps", C.c_int32), ("nsets", C.c_int32), ("periods", C.POINTER(C.c_int32)),
                ("sign_notional", _dp), ("fixed_rate", _dp),
                ("netting_set", C.POINTER(C.c_int32)), ("hazard_rate", _dp),
                ("recovery", C.c_double), ("horizon", C.c_double), ("dtau", C.c_double)]


class _Tuning(C.Structure):
    _fields_ = [("block_size", C.c_int32), ("blocks_per_cu", C.c_int32),
                ("use_graph", C.c_int32), ("fuse_statistics", C.c_int32),
                ("pfe_stage_slots", C.c_int32), ("path_tiles", C.c_int32),
                ("path_waves", C.c_int32), ("value_tiles", C.c_int32),
                ("value_waves", C.c_int32), ("value_stage", C.c_int32)]


class _PfResults(C.Structure):
    _fields_ = [("expected_exposure", _dp), ("pfe", _dp), ("cva", _dp), ("cva_stderr", _dp),
                ("exposures", _dp), ("timings_ms", _dp)]


class _Calibration(C.Structure):
    _fields_ = [("curves", _dp), ("ncurves", C.c_int64), ("window", C.c_int64),
                ("step", C.c_int64), ("nwindows", C.c_int32), ("nfactors", C.c_int32),
                ("obs_per_year", C.c_double), ("dtau", C.c_double),
                ("fit_degrees", C.POINTER(C.c_int32)), ("device", C.c_int32)]


class _CalibResults(C.Structure):
    _fields_ = [("eigenvalues", _dp), ("volatilities", _dp), ("drifts", _dp),
                ("raw_loadings", _dp), ("timings_ms", _dp)]


class _Results(C.Structure):
    _fields_ = [("expected_exposure", _dp), ("pfe", _dp), ("cva", _dp), ("cva_stderr", _dp),
                ("sums", _dp), ("num_x", _dp), ("num_y", _dp), ("exposures", _dp),
                ("timings_ms", _dp)]


_LIB = None


def library() -> C.CDLL:
    """Load libhjm_cva.so (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(
                "libhjm_cva.so not built (%s); run python -m cva_hjm_cuda_amd.build" % LIB_PATH)
        lib = C.CDLL(LIB_PATH)
        _declare(lib)
        _LIB = lib
    return _LIB


def _declare(lib):
    i, d, f = C.c_int, C.c_double, C.c_float
    fp = C.POINTER(C.c_float)
    lib.hjm_status_string.restype = C.c_char_p
    lib.hjm_status_string.argtypes = [i]
    lib.hjm_device_count.argtypes = [C.POINTER(i)]
    lib.hjm_cva_weights.argtypes = [i, d, d, d, _dp]
    lib.hjm_hjm_drift.argtypes = [_dp, d, _dp]
    lib.hjm_plan_create.argtypes = [C.POINTER(C.c_void_p), C.POINTER(_Market), C.POINTER(_Swap),
                                    C.POINTER(_Config)]
    lib.hjm_plan_execute.argtypes = [C.c_void_p, C.POINTER(_Results)]
    lib.hjm_plan_destroy.argtypes = [C.c_void_p]
    lib.hjm_plan_destroy.restype = None
    for name in ("hjm_simulate_reference", "hjm_simulate_reference_blockwise"):
        getattr(lib, name).argtypes = [C.POINTER(_Market), C.POINTER(_Swap), C.POINTER(_Config),
                                       C.POINTER(_Results)]
    lib.hjm_cva_greeks_reference.argtypes = [C.POINTER(_Market), C.POINTER(_Swap),
                                             C.POINTER(_Config), C.POINTER(_Greeks)]
    lib.hjm_simulate_reference_noalloc.argtypes = [C.POINTER(_Market), C.POINTER(_Swap),
                                                   C.POINTER(_Config), C.c_void_p, C.c_void_p]
    lib.hjm_portfolio_plan_create.argtypes = [C.POINTER(C.c_void_p), C.POINTER(_Market),
                                              C.POINTER(_Portfolio), C.POINTER(_Config),
                                              C.POINTER(_Tuning)]
    lib.hjm_portfolio_plan_execute.argtypes = [C.c_void_p, C.POINTER(_PfResults)]
    lib.hjm_portfolio_plan_destroy.argtypes = [C.c_void_p]
    lib.hjm_portfolio_plan_destroy.restype = None
    lib.hjm_simulate_portfolio.argtypes = [C.POINTER(_Market), C.POINTER(_Portfolio),
                                           C.POINTER(_Config), C.POINTER(_Tuning),
                                           C.POINTER(_PfResults)]
    lib.hjm_portfolio_autotune.argtypes = [C.POINTER(_Market), C.POINTER(_Portfolio),
                                           C.POINTER(_Config), C.POINTER(_Tuning),
                                           C.POINTER(C.c_double)]
    lib.hjm_portfolio_cva_greeks.argtypes = [C.POINTER(_Market), C.POINTER(_Portfolio),
                                             C.POINTER(_Config), C.POINTER(_Tuning),
                                             C.POINTER(_Greeks)]
    lib.hjm_calculate_exposure_multi_gpu_f64.argtypes = [_dp, d, d, d, d, _dp, _dp, _dp, _dp, d,
                                                         i, i, d]
    lib.hjm_calculate_exposure_multi_gpu_f32.argtypes = [fp, f, f, f, f, fp, fp, fp, fp, f, i, i,
                                                         f]
    lib.hjm_block_map.argtypes = [_dp, _dp, d, d, C.c_int32, fp]
    lib.hjm_block_map64.argtypes = [_dp, _dp, d, d, C.c_int32, C.c_int32, _dp]
    lib.hjm_select_kth.argtypes = [_dp, C.c_int32, C.c_int64, C.c_int64, C.c_int32, _dp]
    lib.hjm_calibrate_pca.argtypes = [C.POINTER(_Calibration), C.POINTER(_CalibResults)]
    lib.hjm_generate_normals.argtypes = [C.c_uint64, C.c_int64, C.c_int64, C.c_int32, C.c_int32,
                                         C.c_int32, _dp]
    _fp32 = C.POINTER(C.c_float)
    lib.hjm_box_muller_parts.argtypes = [C.POINTER(C.c_uint32), C.c_int64, C.c_int32, _fp32,
                                         _fp32, _fp32]


def _ptr(a: Optional[np.ndarray]):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def _check(code: int, what: str):
    if code != 0:
        raise HJMError(code, what)


# ---- product and market (product.h:4-17) -----------------------------------------
@dataclass
class InterestRateSwap:
    """The fields of the reference's InterestRateSwap<real> that its path reads."""
    notional: float
    K: float
    expiry: float
    dtau: float = 0.5
    pricing_points: Optional[np.ndarray] = None
    floating_schedule: Optional[np.ndarray] = None
    fixed_schedule: Optional[np.ndarray] = None


@dataclass
class Market:
    spot_rates: np.ndarray
    drifts: np.ndarray
    volatilities: np.ndarray
    accrual: np.ndarray

    def __post_init__(self):
        for name, n in (("spot_rates", TENORS), ("drifts", TENORS),
                        ("volatilities", FACTORS * TENORS), ("accrual", TENORS)):
            a = np.array(getattr(self, name), dtype=np.float64, copy=True).reshape(-1)
            if a.size != n:
                raise ValueError("%s must have %d entries" % (name, n))
            setattr(self, name, a)

    def _c(self):
        return _Market(_ptr(self.spot_rates), _ptr(self.drifts), _ptr(self.volatilities),
                       _ptr(self.accrual))


def hjm_drift(volatilities, dtau: float = 0.5) -> np.ndarray:
    """HJM drift of the loadings by the cumulative Simpson rule (library routine)."""
    v = np.ascontiguousarray(np.asarray(volatilities, np.float64).reshape(-1))
    out = np.zeros(TENORS)
    _check(library().hjm_hjm_drift(_ptr(v), dtau, _ptr(out)), "hjm_hjm_drift")
    return out


def cva_weights(hazard: float, recovery: float, ndates: int = TENORS, date_dt: float = 0.5):
    out = np.zeros(ndates)
    _check(library().hjm_cva_weights(ndates, date_dt, hazard, recovery, _ptr(out)),
           "hjm_cva_weights")
    return out


def benchmark_market(dtau: float = 0.5, accrual: float = 0.5) -> Market:
    """The benchmark market: Nelson-Siegel curve (b0=3%, b1=-2%, b2=1%, lambda=2y), the
    three parametric loadings, their Simpson HJM drift, flat accrual."""
    tau = np.arange(TENORS) * dtau
    x = tau / 2.0
    spot = 0.03 - 0.02 * np.exp(-x) + 0.01 * x * np.exp(-x)
    vols = np.concatenate([np.full(TENORS, 0.0064),
                           -0.0036 + 0.0006 * tau - 2e-5 * tau ** 2,
                           -0.0047 + 0.0018 * tau - 1.5e-4 * tau ** 2 + 3e-6 * tau ** 3])
    return Market(spot, hjm_drift(vols, dtau), vols, np.full(TENORS, accrual))


def device_count() -> int:
    n = C.c_int(0)
    _check(library().hjm_device_count(C.byref(n)), "hjm_device_count")
    return n.value


# ---- reference-semantics simulation ------------------------------------------------
@dataclass
class ReferenceResult:
    expected_exposure: np.ndarray
    pfe: np.ndarray
    cva: float
    cva_stderr: float
    sums: np.ndarray
    timings_ms: np.ndarray
    num_x: Optional[np.ndarray] = None
    num_y: Optional[np.ndarray] = None
    exposures: Optional[np.ndarray] = None


KERNEL_AUTO, KERNEL_VECTOR, KERNEL_MATRIX = 0, 1, 2


def _config(paths, dt, precision, rng, seed, normals, hazard, recovery, pfe_quantile, device,
            path_offset, path_kernel=KERNEL_AUTO, regenerate_normals=False):
    return _Config(int(paths), int(path_offset), float(dt), int(precision), int(rng), int(seed),
                   _ptr(normals), float(hazard), float(recovery), float(pfe_quantile),
                   int(device), int(path_kernel), int(bool(regenerate_normals)))


def _host_normals(normals, paths: int, horizon: float, dt: float, what: str) -> np.ndarray:
    """Caller-supplied normals as the C ABI reads them: contiguous float64
    [paths][round(horizon / dt)][3]. Any other shape is rejected here, since the library
    copies exactly paths x steps x 3 values from the pointer."""
    z = np.ascontiguousarray(normals, np.float64)
    want = (int(paths), int(round(horizon / dt)), 3)
    if z.shape != want:
        raise ValueError("%s: normals must have shape (paths, horizon/dt, 3) = %s, got %s"
                         % (what, want, z.shape))
    return z


def simulate_reference(market: Market, swap: InterestRateSwap, paths: int, dt: float = 0.01,
                       precision: int = FP64, seed: int = 1234,
                       normals: Optional[np.ndarray] = None, hazard: float = 0.02,
                       recovery: float = 0.4, pfe_quantile: float = 0.975, device: int = 0,
                       path_offset: int = 0, keep_paths: bool = False,
                       blockwise: bool = False, path_kernel: int = KERNEL_AUTO) -> ReferenceResult:
    """Simulate the reference semantics on the GPU and return EE, PFE, CVA (+ SE).

    ``normals`` (float64, [paths][expiry/dt][3]) switches to caller-supplied variates;
    otherwise the in-kernel Philox generator with ``seed`` is used. ``blockwise`` runs the
    reference's own kernel decomposition (the baseline form) instead.
    """
    lib = library()
    rng = RNG_PHILOX
    if normals is not None:
        normals = _host_normals(normals, paths, swap.expiry, dt, "simulate_reference")
        rng = RNG_HOST
    mk = market._c()
    sw = _Swap(swap.notional, swap.K, swap.expiry, swap.dtau)
    cfg = _config(paths, dt, precision, rng, seed, normals, hazard, recovery, pfe_quantile,
                  device, path_offset, path_kernel)
    ee, pfe = np.zeros(TENORS), np.zeros(TENORS)
    cva, se = np.zeros(1), np.zeros(1)
    sums, tm = np.zeros(STAT_SLOTS), np.zeros(5)
    nx = ny = ex = None
    if keep_paths:
        nx, ny, ex = (np.zeros((paths, TENORS)) for _ in range(3))
    res = _Results(_ptr(ee), _ptr(pfe), _ptr(cva), _ptr(se), _ptr(sums), _ptr(nx), _ptr(ny),
                   _ptr(ex), _ptr(tm))
    fn = lib.hjm_simulate_reference_blockwise if blockwise else lib.hjm_simulate_reference
    _check(fn(C.byref(mk), C.byref(sw), C.byref(cfg), C.byref(res)), fn.__name__)
    return ReferenceResult(ee, pfe, float(cva[0]), float(se[0]), sums, tm, nx, ny, ex)


def simulate_reference_noalloc(market: Market, swap: InterestRateSwap, paths: int,
                               d_sums: int, stream: int = 0, dt: float = 0.01,
                               precision: int = FP64, seed: int = 1234, hazard: float = 0.02,
                               recovery: float = 0.4, device: int = 0, path_offset: int = 0,
                               path_kernel: int = KERNEL_AUTO) -> None:
    """Enqueue hjm_simulate_reference_noalloc on ``stream`` (a hipStream_t as an int, 0 =
    the null stream): the 53 raw sums of the run go to the device address ``d_sums``; the
    library allocates nothing and does not wait for the work."""
    mk = market._c()
    sw = _Swap(swap.notional, swap.K, swap.expiry, swap.dtau)
    cfg = _config(paths, dt, precision, RNG_PHILOX, seed, None, hazard, recovery, 0.0, device,
                  path_offset, path_kernel)
    _check(library().hjm_simulate_reference_noalloc(C.byref(mk), C.byref(sw), C.byref(cfg),
                                                    C.c_void_p(d_sums), C.c_void_p(stream)),
           "hjm_simulate_reference_noalloc")


@dataclass
class Greeks:
    cva: float
    cva_stderr: float
    d_spot: np.ndarray       # dCVA / d f(0, tau_j)                      [51]
    d_vol: np.ndarray        # dCVA / d sigma_kj, drift table held fixed [153]
    d_drift: np.ndarray      # dCVA / d m_j                              [51]
    d_vol_total: np.ndarray  # d_vol + (dm/dsigma)^T d_drift (Simpson HJM drift)  [153]
    timings_ms: np.ndarray   # forward, adjoint, total (device)
    stored_normals_bytes: int = 0  # shocks kept by the forward for the sweep (0: regenerated)


def cva_greeks(market: Market, swap: InterestRateSwap, paths: int, dt: float = 0.01,
               precision: int = FP32_KAHAN, seed: int = 1234,
               normals: Optional[np.ndarray] = None, hazard: float = 0.02,
               recovery: float = 0.4, device: int = 0, path_offset: int = 0,
               path_kernel: int = KERNEL_AUTO, regenerate_normals: bool = False) -> Greeks:
    """CVA and its 51 + 153 (+ 51 drift) pathwise-adjoint sensitivities on the GPU.
    ``regenerate_normals``: the reverse sweep regenerates the shocks instead of reading the
    matrix-core forward's stored copy (no paths x steps x 3 FP32 buffer)."""
    rng = RNG_PHILOX
    if normals is not None:
        normals = _host_normals(normals, paths, swap.expiry, dt, "cva_greeks")
        rng = RNG_HOST
    mk = market._c()
    sw = _Swap(swap.notional, swap.K, swap.expiry, swap.dtau)
    cfg = _config(paths, dt, precision, rng, seed, normals, hazard, recovery, 0.0, device,
                  path_offset, path_kernel, regenerate_normals)
    cva, se = np.zeros(1), np.zeros(1)
    ds, dv, dm, dvt, tm = np.zeros(TENORS), np.zeros(3 * TENORS), np.zeros(TENORS), \
        np.zeros(3 * TENORS), np.zeros(3)
    nb = np.zeros(1, np.int64)
    g = _Greeks(_ptr(cva), _ptr(se), _ptr(ds), _ptr(dv), _ptr(dm), _ptr(dvt), _ptr(tm),
                nb.ctypes.data_as(C.POINTER(C.c_int64)))
    _check(library().hjm_cva_greeks_reference(C.byref(mk), C.byref(sw), C.byref(cfg),
                                              C.byref(g)), "hjm_cva_greeks_reference")
    return Greeks(float(cva[0]), float(se[0]), ds, dv, dm, dvt, tm, int(nb[0]))


class ReferencePlan:
    """A reusable plan: device buffers and stream held across executions (benchmarks)."""

    def __init__(self, market: Market, swap: InterestRateSwap, paths: int, dt: float = 0.01,
                 precision: int = FP64, seed: int = 1234, hazard: float = 0.02,
                 recovery: float = 0.4, pfe_quantile: float = 0.975, device: int = 0,
                 path_kernel: int = KERNEL_AUTO):
        self._lib = library()
        self._market = market
        mk = market._c()
        sw = _Swap(swap.notional, swap.K, swap.expiry, swap.dtau)
        cfg = _config(paths, dt, precision, RNG_PHILOX, seed, None, hazard, recovery,
                      pfe_quantile, device, 0, path_kernel)
        self._h = C.c_void_p()
        _check(self._lib.hjm_plan_create(C.byref(self._h), C.byref(mk), C.byref(sw),
                                         C.byref(cfg)), "hjm_plan_create")
        self.paths = paths
        self.ee, self.pfe = np.zeros(TENORS), np.zeros(TENORS)
        self.cva, self.se = np.zeros(1), np.zeros(1)
        self.sums, self.timings_ms = np.zeros(STAT_SLOTS), np.zeros(5)
        self._res = _Results(_ptr(self.ee), _ptr(self.pfe), _ptr(self.cva), _ptr(self.se),
                             _ptr(self.sums), None, None, None, _ptr(self.timings_ms))

    def execute(self):
        _check(self._lib.hjm_plan_execute(self._h, C.byref(self._res)), "hjm_plan_execute")
        return self

    def close(self):
        if self._h:
            self._lib.hjm_plan_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def block_map(market: Market, dt: float = 0.01, dtau: float = 0.5,
              portfolio: bool = False) -> np.ndarray:
    """A matrix-core path kernel's affine 50-step block map, float32 with 224 columns in
    the kernel's K order: [102][224] for the reference semantics, [52][224] for the
    portfolio semantics."""
    out = np.zeros((52 if portfolio else 102, 224), np.float32)
    _check(library().hjm_block_map(_ptr(market.drifts), _ptr(market.volatilities), dt, dtau,
                                   int(bool(portfolio)),
                                   out.ctypes.data_as(C.POINTER(C.c_float))), "hjm_block_map")
    return out


def block_map64(market: Market, dt: float = 0.01, dtau: float = 0.5,
                portfolio: bool = False, mtiles: int = 0) -> np.ndarray:
    """The FP64 matrix-core kernels' block map as they read it: A operands of
    v_mfma_f64_16x16x4f64, [M-tile][52 instructions][64 lanes] (hjm_block_map64); mtiles 0
    means all (7 reference, 4 portfolio)."""
    mt = mtiles or (4 if portfolio else 7)
    out = np.zeros((mt, 52, 64))
    _check(library().hjm_block_map64(_ptr(market.drifts), _ptr(market.volatilities), dt, dtau,
                                     int(bool(portfolio)), int(mt), _ptr(out)), "hjm_block_map64")
    return out


def select_kth(values, k: int, device: int = 0) -> np.ndarray:
    """k-th smallest (1-based) of each row of non-negative values [rows][n] on the GPU
    (the exact radix selection behind the PFE)."""
    v = np.ascontiguousarray(np.atleast_2d(values), np.float64)
    out = np.zeros(v.shape[0])
    _check(library().hjm_select_kth(_ptr(v), v.shape[0], v.shape[1], int(k), int(device),
                                    _ptr(out)), "hjm_select_kth")
    return out


def generate_normals(seed: int, paths: int, nsteps: int, path_offset: int = 0,
                     device: int = 0, stride: int = 50) -> np.ndarray:
    """The in-kernel generator's normals, [paths][nsteps][3], computed on the GPU, for runs
    whose exposure dates are `stride` steps apart (dtau / dt)."""
    out = np.zeros((paths, nsteps, 3))
    _check(library().hjm_generate_normals(seed, path_offset, paths, nsteps, stride, device,
                                          _ptr(out)),
           "hjm_generate_normals")
    return out


def box_muller_parts(words: np.ndarray, device: int = 0):
    """The generator's Box-Muller factors of 32-bit Philox words, on the GPU: (radius of
    each word as the first of a pair, cosine and sine of each as the second), float32."""
    w = np.ascontiguousarray(words, np.uint32)
    out = [np.zeros(w.size, np.float32) for _ in range(3)]
    fp = [o.ctypes.data_as(C.POINTER(C.c_float)) for o in out]
    _check(library().hjm_box_muller_parts(w.ctypes.data_as(C.POINTER(C.c_uint32)), w.size,
                                          int(device), *fp), "hjm_box_muller_parts")
    return tuple(out)


# ---- drop-in for calculateExposureMultiGPU (simulation_gpu.h:10,12) ---------------
def calculate_exposure_multi_gpu(expected_exposure: np.ndarray, payOff: InterestRateSwap,
                                 accrual, spot_rates, drift, volatilities, scale: float,
                                 num_gpus: int, scenarios: int, dt: float) -> None:
    """Fill ``expected_exposure`` (51 entries, float64 or float32) like the reference.

    float64 output runs the FP64 path, float32 output the FP32 + compensated path; the
    scenarios are split into ``num_gpus`` contiguous shards (devices g % device_count),
    normals come from the in-kernel generator with seed 1234; ``scale`` is unused.
    """
    lib = library()
    ee = expected_exposure
    if ee.dtype == np.float64:
        arrs = [np.ascontiguousarray(a, np.float64) for a in (accrual, spot_rates, drift,
                                                              volatilities)]
        code = lib.hjm_calculate_exposure_multi_gpu_f64(
            ee.ctypes.data_as(_dp), payOff.notional, payOff.K, payOff.expiry, payOff.dtau,
            *[_ptr(a) for a in arrs], float(scale), int(num_gpus), int(scenarios), float(dt))
    elif ee.dtype == np.float32:
        fp = C.POINTER(C.c_float)
        arrs = [np.ascontiguousarray(a, np.float32) for a in (accrual, spot_rates, drift,
                                                              volatilities)]
        code = lib.hjm_calculate_exposure_multi_gpu_f32(
            ee.ctypes.data_as(fp), payOff.notional, payOff.K, payOff.expiry, payOff.dtau,
            *[a.ctypes.data_as(fp) for a in arrs], float(scale), int(num_gpus),
            int(scenarios), float(dt))
    else:
        raise TypeError("expected_exposure must be float64 or float32")
    _check(code, "calculateExposureMultiGPU")


# ---- portfolio semantics ------------------------------------------------------------
@dataclass
class Tuning:
    """Launch shapes of a portfolio plan (0: the kernel's default; see hjm_tuning)."""
    block_size: int = 256          # VALU path kernel
    blocks_per_cu: int = 0         # persistent path kernel (matrix-core: 0 or 1)
    use_graph: bool = False
    fuse_statistics: bool = False  # FP32: valuation and statistics in one pass
    pfe_stage_slots: int = 0       # fused pass: 0 = 32
    path_tiles: int = 0            # matrix-core path kernel: N-tiles per wave (0 = 2)
    path_waves: int = 0            # ... waves per workgroup (0 = 8)
    value_tiles: int = 0           # matrix-core valuation: N-tiles per wave (0 = 4)
    value_waves: int = 0           # ... waves per workgroup (0 = 4)
    value_stage: int = 0           # ... M-tiles staged in LDS at a time (0 = 8)

    def _c(self):
        return _Tuning(self.block_size, self.blocks_per_cu, int(self.use_graph),
                       int(self.fuse_statistics), int(self.pfe_stage_slots), self.path_tiles,
                       self.path_waves, self.value_tiles, self.value_waves, self.value_stage)

    @staticmethod
    def _from_c(t):
        return Tuning(t.block_size, t.blocks_per_cu, bool(t.use_graph), bool(t.fuse_statistics),
                      t.pfe_stage_slots, t.path_tiles, t.path_waves, t.value_tiles,
                      t.value_waves, t.value_stage)

    def as_dict(self):
        return {k: (int(v) if isinstance(v, bool) else v) for k, v in self.__dict__.items()}


@dataclass
class PortfolioResult:
    expected_exposure: np.ndarray   # [nsets][ndates]
    pfe: np.ndarray                 # [nsets][ndates]
    cva: np.ndarray                 # [nsets]
    cva_stderr: np.ndarray          # [nsets]
    timings_ms: np.ndarray          # paths, valuation, statistics, PFE, total
    exposures: Optional[np.ndarray] = None  # [paths][nsets][ndates]


def _pf_c(pf, horizon, recovery):
    return _Portfolio(pf.nswaps, pf.nsets, pf.periods.ctypes.data_as(C.POINTER(C.c_int32)),
                      _ptr(pf.sign_notional), _ptr(pf.fixed_rate),
                      pf.set_of.ctypes.data_as(C.POINTER(C.c_int32)), _ptr(pf.hazard),
                      float(recovery), float(horizon), float(pf.dtau))


class PortfolioPlan:
    """Device buffers, coefficient tables and (optionally) a captured hipGraph for
    repeated runs of one portfolio configuration."""

    def __init__(self, market: Market, portfolio, paths: int, horizon: float = 30.0,
                 dt: float = 0.01, precision: int = FP32_KAHAN, seed: int = 2024,
                 normals: Optional[np.ndarray] = None, recovery: float = 0.4,
                 pfe_quantile: float = 0.975, tuning: Optional[Tuning] = None, device: int = 0,
                 path_offset: int = 0, path_kernel: int = KERNEL_AUTO):
        self._lib = library()
        self._keep = (market, portfolio, normals)
        rng = RNG_PHILOX
        if normals is not None:
            normals = _host_normals(normals, paths, horizon, dt, "PortfolioPlan")
            self._keep = (market, portfolio, normals)
            rng = RNG_HOST
        mk = market._c()
        pfc = _pf_c(portfolio, horizon, recovery)
        cfg = _config(paths, dt, precision, rng, seed, normals, 0.0, recovery, pfe_quantile,
                      device, path_offset, path_kernel)
        tun = tuning._c() if tuning is not None else None
        self._h = C.c_void_p()
        _check(self._lib.hjm_portfolio_plan_create(C.byref(self._h), C.byref(mk), C.byref(pfc),
                                                   C.byref(cfg),
                                                   C.byref(tun) if tun is not None else None),
               "hjm_portfolio_plan_create")
        self.paths, self.nsets = paths, portfolio.nsets
        self.ndates = int(round(horizon / portfolio.dtau)) + 1

    def execute(self, keep_paths: bool = False) -> PortfolioResult:
        G, D = self.nsets, self.ndates
        ee, pfe = np.zeros((G, D)), np.zeros((G, D))
        cva, se, tm = np.zeros(G), np.zeros(G), np.zeros(5)
        ex = np.zeros((self.paths, G, D)) if keep_paths else None
        r = _PfResults(_ptr(ee), _ptr(pfe), _ptr(cva), _ptr(se), _ptr(ex), _ptr(tm))
        _check(self._lib.hjm_portfolio_plan_execute(self._h, C.byref(r)),
               "hjm_portfolio_plan_execute")
        return PortfolioResult(ee, pfe, cva, se, tm, ex)

    def close(self):
        if self._h:
            self._lib.hjm_portfolio_plan_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def simulate_portfolio(market: Market, portfolio, paths: int, keep_paths: bool = False,
                       **kw) -> PortfolioResult:
    plan = PortfolioPlan(market, portfolio, paths, **kw)
    try:
        return plan.execute(keep_paths)
    finally:
        plan.close()


def autotune_portfolio(market: Market, portfolio, paths: int, horizon: float = 30.0,
                       dt: float = 0.01, precision: int = FP32_KAHAN, seed: int = 2024,
                       recovery: float = 0.4, pfe_quantile: float = 0.975, device: int = 0,
                       path_kernel: int = KERNEL_AUTO):
    """Fastest launch shapes for this problem, measured on the device: returns (Tuning
    with use_graph set, eager ms of one execute with it). Matrix-core FP32 plans tune the
    path kernel (N-tiles, waves, persistent grid), the valuation (N-tiles, waves, LDS
    stage) and the fused against the exposure pass; VALU path kernels their block size
    and persistent grid; FP64 plans on the matrix cores keep the defaults
    (hjm_portfolio_autotune)."""
    mk = market._c()
    pfc = _pf_c(portfolio, horizon, recovery)
    cfg = _config(paths, dt, precision, RNG_PHILOX, seed, None, 0.0, recovery, pfe_quantile,
                  device, 0, path_kernel)
    t = _Tuning()
    ms = C.c_double()
    _check(library().hjm_portfolio_autotune(C.byref(mk), C.byref(pfc), C.byref(cfg), C.byref(t),
                                            C.byref(ms)), "hjm_portfolio_autotune")
    return Tuning._from_c(t), ms.value


def portfolio_cva_greeks(market: Market, portfolio, paths: int, horizon: float = 30.0,
                         dt: float = 0.01, precision: int = FP32_KAHAN, seed: int = 2024,
                         normals: Optional[np.ndarray] = None, recovery: float = 0.4,
                         tuning: Optional[Tuning] = None, device: int = 0,
                         path_kernel: int = KERNEL_AUTO, regenerate_normals: bool = False,
                         path_offset: int = 0) -> Greeks:
    """Total portfolio CVA and its 51 + 153 (+ 51 drift) pathwise-adjoint sensitivities.
    ``regenerate_normals`` as for cva_greeks."""
    rng = RNG_PHILOX
    if normals is not None:
        normals = _host_normals(normals, paths, horizon, dt, "portfolio_cva_greeks")
        rng = RNG_HOST
    mk = market._c()
    pfc = _pf_c(portfolio, horizon, recovery)
    cfg = _config(paths, dt, precision, rng, seed, normals, 0.0, recovery, 0.0, device,
                  path_offset, path_kernel, regenerate_normals)
    tun = tuning._c() if tuning is not None else None
    cva, se = np.zeros(1), np.zeros(1)
    ds, dv, dm, dvt, tm = (np.zeros(TENORS), np.zeros(3 * TENORS), np.zeros(TENORS),
                           np.zeros(3 * TENORS), np.zeros(3))
    nb = np.zeros(1, np.int64)
    g = _Greeks(_ptr(cva), _ptr(se), _ptr(ds), _ptr(dv), _ptr(dm), _ptr(dvt), _ptr(tm),
                nb.ctypes.data_as(C.POINTER(C.c_int64)))
    _check(library().hjm_portfolio_cva_greeks(C.byref(mk), C.byref(pfc), C.byref(cfg),
                                              C.byref(tun) if tun is not None else None,
                                              C.byref(g)), "hjm_portfolio_cva_greeks")
    return Greeks(float(cva[0]), float(se[0]), ds, dv, dm, dvt, tm, int(nb[0]))


# ---- calibration upstream of the path -------------------------------------------------
@dataclass
class Calibration:
    eigenvalues: np.ndarray    # [windows][51] annualised covariance eigenvalues, descending
    volatilities: np.ndarray   # [windows][153] loadings (factor-major), fitted where asked
    drifts: np.ndarray         # [windows][51] HJM drift of the loadings
    raw_loadings: np.ndarray   # [windows][153] sqrt(lambda_k) v_k before fitting
    timings_ms: np.ndarray     # [1] calibration kernel

    def market(self, spot_rates, accrual, window: int = -1) -> Market:
        """A Market from one window's loadings and drift and the given curve/accrual."""
        return Market(np.asarray(spot_rates, np.float64), self.drifts[window],
                      self.volatilities[window], np.asarray(accrual, np.float64))


def calibrate_pca(curves, window: Optional[int] = None, step: int = 1, nwindows: int = 1,
                  obs_per_year: float = 252.0, fit_degrees=None, dtau: float = 0.5,
                  device: int = 0) -> Calibration:
    """Rolling-window PCA calibration of the three HJM loadings from historical forward
    curves [ncurves][51] (oldest first) on the GPU; ``window`` defaults to all curves.
    ``fit_degrees`` (3 ints, < 0 = raw) smooths each loading by a least-squares polynomial
    in the tenor, e.g. (0, 2, 3) for the benchmark market's parametric forms."""
    c = np.ascontiguousarray(curves, np.float64)
    if c.ndim != 2 or c.shape[1] != TENORS:
        raise ValueError("curves must have shape (ncurves, 51)")
    window = c.shape[0] if window is None else int(window)
    deg = None
    if fit_degrees is not None:
        deg = np.ascontiguousarray(fit_degrees, np.int32)
        if deg.shape != (3,):
            raise ValueError("fit_degrees needs one degree per factor (3)")
    W = int(nwindows)
    ev, vols, dr, raw, tm = (np.zeros((W, TENORS)), np.zeros((W, 3 * TENORS)),
                             np.zeros((W, TENORS)), np.zeros((W, 3 * TENORS)), np.zeros(1))
    cal = _Calibration(_ptr(c), c.shape[0], window, int(step), W, 3, float(obs_per_year),
                       float(dtau),
                       deg.ctypes.data_as(C.POINTER(C.c_int32)) if deg is not None else None,
                       int(device))
    res = _CalibResults(_ptr(ev), _ptr(vols), _ptr(dr), _ptr(raw), _ptr(tm))
    _check(library().hjm_calibrate_pca(C.byref(cal), C.byref(res)), "hjm_calibrate_pca")
    return Calibration(ev, vols, dr, raw, tm)
