"""ctypes binding of the CPU oracle (oracle/_build/libhjm_oracle.so) for the tests.

The oracle is test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg load it. It is built by ``make -C oracle`` (also run by
__graft_entry__.build()).
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "_build", "libhjm_oracle.so")
TENORS = 51

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")


class Oracle:
    def __init__(self, lib):
        self.lib = lib
        L = lib
        L.orc_philox4x32_10.argtypes = [_u32p, _u32p, _u32p]
        L.orc_box_muller.argtypes = [C.c_uint32, C.c_uint32, C.POINTER(C.c_float),
                                     C.POINTER(C.c_float)]
        L.orc_stream_call.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32,
                                      C.c_uint32, _f32p]
        L.orc_step_normals.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, _f32p]
        L.orc_fill_normals.argtypes = [C.c_uint64, C.c_int64, C.c_int, C.c_int, C.c_int, _dp]
        L.orc_cumulative_simpson.argtypes = [_dp, C.c_int, C.c_double, _dp]
        L.orc_hjm_drift.argtypes = [_dp, C.c_int, C.c_int, C.c_double, _dp]
        L.orc_nelson_siegel.argtypes = [C.c_double] * 4 + [C.c_int, C.c_double, _dp]
        L.orc_poly_vols.argtypes = [C.c_int, C.c_double, _dp]
        L.orc_reference_paths.argtypes = [_dp, _dp, _dp, C.c_double, C.c_double, C.c_double,
                                          C.c_int, _dp, _dp, _dp]
        L.orc_reference_paths.restype = C.c_int
        L.orc_reference_exposure.argtypes = [_dp, _dp, _dp, C.c_double, C.c_double, C.c_double,
                                             C.c_int, _dp]
        L.orc_expected_exposure.argtypes = [_dp, C.c_int, C.c_int, _dp]
        L.orc_cva_weights.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, _dp]
        L.orc_quantile_rank.argtypes = [C.c_int64, C.c_double]
        L.orc_quantile_rank.restype = C.c_int64
        L.orc_stats.argtypes = [_dp, C.c_int, C.c_int, _dp, C.c_double, _dp, _dp,
                                C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_reference_cva_greeks.argtypes = [_dp, _dp, _dp, _dp, C.c_double, C.c_double,
                                               C.c_double, C.c_double, C.c_double, C.c_int, _dp,
                                               _dp, C.POINTER(C.c_double), _dp, _dp, _dp]
        L.orc_reference_cva_greeks.restype = C.c_int
        L.orc_drift_vjp.argtypes = [_dp, C.c_double, _dp, _dp]
        _ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
        L.orc_portfolio_exposure.argtypes = [_dp, _dp, _dp, C.c_double, C.c_double, C.c_double,
                                             C.c_int, _dp, C.c_int, _ip, _dp, _dp, _ip, C.c_int,
                                             _dp, C.c_void_p, C.c_void_p]
        L.orc_portfolio_exposure.restype = C.c_int
        L.orc_portfolio_cva_greeks.argtypes = [_dp, _dp, _dp, C.c_double, C.c_double,
                                               C.c_double, C.c_int, _dp, C.c_int, _ip, _dp,
                                               _dp, _ip, C.c_int, _dp, C.POINTER(C.c_double),
                                               _dp, _dp, _dp]
        L.orc_portfolio_cva_greeks.restype = C.c_int
        L.orc_time_reference_paths.argtypes = [_dp, _dp, _dp, C.c_double, C.c_double, C.c_double,
                                               C.c_int, _dp, _dp, _dp, C.c_int]
        L.orc_time_reference_paths.restype = C.c_double

    # -- rng ---------------------------------------------------------------------
    def philox(self, ctr, key):
        out = np.zeros(4, np.uint32)
        self.lib.orc_philox4x32_10(np.asarray(ctr, np.uint32), np.asarray(key, np.uint32), out)
        return out

    def box_muller(self, a, b):
        z0, z1 = C.c_float(), C.c_float()
        self.lib.orc_box_muller(int(a), int(b), C.byref(z0), C.byref(z1))
        return z0.value, z1.value

    def stream_call(self, seed, path, block, cls, call):
        z = np.zeros(4, np.float32)
        self.lib.orc_stream_call(seed, path, block, cls, call, z)
        return z

    def step_normals(self, seed, path, step, stride=50):
        z = np.zeros(3, np.float32)
        self.lib.orc_step_normals(seed, path, step, stride, z)
        return z

    def normals(self, seed, npaths, nsteps, path0=0, stride=50):
        """[npaths][nsteps][3] shocks of a run with date blocks of `stride` steps."""
        out = np.zeros((npaths, nsteps, 3), np.float64)
        self.lib.orc_fill_normals(seed, path0, npaths, nsteps, stride, out)
        return out

    # -- market ------------------------------------------------------------------
    def cumulative_simpson(self, y, h):
        y = np.ascontiguousarray(y, np.float64)
        out = np.zeros_like(y)
        self.lib.orc_cumulative_simpson(y, len(y), h, out)
        return out

    def hjm_drift(self, vols, h=0.5):
        vols = np.ascontiguousarray(vols, np.float64).reshape(-1)
        nt = TENORS
        out = np.zeros(nt)
        self.lib.orc_hjm_drift(vols, len(vols) // nt, nt, h, out)
        return out

    def nelson_siegel(self, b0=0.03, b1=-0.02, b2=0.01, lam=2.0, h=0.5):
        out = np.zeros(TENORS)
        self.lib.orc_nelson_siegel(b0, b1, b2, lam, TENORS, h, out)
        return out

    def poly_vols(self, h=0.5):
        out = np.zeros(3 * TENORS)
        self.lib.orc_poly_vols(TENORS, h, out)
        return out

    # -- reference semantics ------------------------------------------------------
    def reference_paths(self, spot, drift, vols, expiry, dtau, dt, normals):
        P = normals.shape[0]
        nx = np.zeros((P, TENORS))
        ny = np.zeros((P, TENORS))
        rc = self.lib.orc_reference_paths(_c(spot), _c(drift), _c(vols), expiry, dtau, dt, P,
                                          _c(normals), nx, ny)
        if rc:
            raise ValueError("orc_reference_paths rc=%d" % rc)
        return nx, ny

    def reference_exposure(self, nx, ny, accrual, notional, K, dtau):
        P = nx.shape[0]
        ex = np.zeros((P, TENORS))
        self.lib.orc_reference_exposure(_c(nx), _c(ny), _c(accrual), notional, K, dtau, P, ex)
        return ex

    def expected_exposure(self, exposures):
        P = exposures.shape[0]
        ee = np.zeros(exposures.shape[1])
        self.lib.orc_expected_exposure(_c(exposures), P, exposures.shape[1], ee)
        return ee

    def cva_weights(self, ndates=TENORS, date_dt=0.5, hazard=0.02, recovery=0.4):
        w = np.zeros(ndates)
        self.lib.orc_cva_weights(ndates, date_dt, hazard, recovery, w)
        return w

    def quantile_rank(self, n, q):
        return self.lib.orc_quantile_rank(n, q)

    def stats(self, exposures, w, q=0.975):
        P, D = exposures.shape
        ee, pfe = np.zeros(D), np.zeros(D)
        cva, se = C.c_double(), C.c_double()
        self.lib.orc_stats(_c(exposures), P, D, _c(w), q, ee, pfe, C.byref(cva), C.byref(se))
        return ee, pfe, cva.value, se.value

    def reference_cva(self, spot, drift, vols, accrual, notional, K, expiry, dtau, dt, normals,
                      w):
        nx, ny = self.reference_paths(spot, drift, vols, expiry, dtau, dt, normals)
        ex = self.reference_exposure(nx, ny, accrual, notional, K, dtau)
        return float(np.mean(ex @ w))

    def reference_cva_greeks(self, spot, drift, vols, accrual, notional, K, expiry, dtau, dt,
                             normals, w):
        P = normals.shape[0]
        cva = C.c_double()
        ds, dv, dm = np.zeros(TENORS), np.zeros(3 * TENORS), np.zeros(TENORS)
        rc = self.lib.orc_reference_cva_greeks(_c(spot), _c(drift), _c(vols), _c(accrual),
                                               notional, K, expiry, dtau, dt, P, _c(normals),
                                               _c(w), C.byref(cva), ds, dv, dm)
        if rc:
            raise ValueError("orc_reference_cva_greeks rc=%d" % rc)
        return cva.value, ds, dv, dm

    def drift_vjp(self, vols, g, h=0.5):
        out = np.zeros(3 * TENORS)
        self.lib.orc_drift_vjp(_c(vols), h, _c(g), out)
        return out

    def portfolio_exposure(self, spot, drift, vols, horizon, dtau, dt, normals, pf,
                           want_curves=False):
        """Exposures [path][set][date] (+ curves [path][date][51], log bank [path][date])."""
        P = normals.shape[0]
        nd = int(round(horizon / dtau)) + 1
        ex = np.zeros((P, pf.nsets, nd))
        cv = np.zeros((P, nd, TENORS)) if want_curves else None
        lb = np.zeros((P, nd)) if want_curves else None
        rc = self.lib.orc_portfolio_exposure(
            _c(spot), _c(drift), _c(vols), horizon, dtau, dt, P, _c(normals), pf.nswaps,
            pf.periods, pf.sign_notional, pf.fixed_rate, pf.set_of, pf.nsets, ex,
            cv.ctypes.data if cv is not None else None, lb.ctypes.data if lb is not None else None)
        if rc:
            raise ValueError("orc_portfolio_exposure rc=%d" % rc)
        return (ex, cv, lb) if want_curves else ex

    def portfolio_weights(self, pf, horizon, recovery=0.4):
        nd = int(round(horizon / pf.dtau)) + 1
        return np.stack([self.cva_weights(nd, pf.dtau, h, recovery) for h in pf.hazard])

    def portfolio_cva(self, spot, drift, vols, horizon, dtau, dt, normals, pf, w):
        ex = self.portfolio_exposure(spot, drift, vols, horizon, dtau, dt, normals, pf)
        return float(np.einsum("pge,ge->", ex, w) / ex.shape[0])

    def portfolio_cva_greeks(self, spot, drift, vols, horizon, dtau, dt, normals, pf, w):
        P = normals.shape[0]
        cva = C.c_double()
        ds, dv, dm = np.zeros(TENORS), np.zeros(3 * TENORS), np.zeros(TENORS)
        rc = self.lib.orc_portfolio_cva_greeks(
            _c(spot), _c(drift), _c(vols), horizon, dtau, dt, P, _c(normals), pf.nswaps,
            pf.periods, pf.sign_notional, pf.fixed_rate, pf.set_of, pf.nsets, _c(w),
            C.byref(cva), ds, dv, dm)
        if rc:
            raise ValueError("orc_portfolio_cva_greeks rc=%d" % rc)
        return cva.value, ds, dv, dm

    def time_reference_paths(self, spot, drift, vols, expiry, dtau, dt, normals, nthreads=1):
        P = normals.shape[0]
        nx = np.zeros((P, TENORS))
        ny = np.zeros((P, TENORS))
        ms = self.lib.orc_time_reference_paths(_c(spot), _c(drift), _c(vols), expiry, dtau, dt,
                                               P, _c(normals), nx, ny, nthreads)
        return ms, nx, ny


def _c(a):
    return np.ascontiguousarray(a, np.float64)


def build():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)


def load(auto_build=True):
    if not os.path.exists(LIB) and auto_build:
        build()
    return Oracle(C.CDLL(LIB))
