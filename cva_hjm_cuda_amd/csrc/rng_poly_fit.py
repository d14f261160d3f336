"""Derive the FP32 polynomial coefficients used by the Box-Muller transform.

The normal generator must give bit-identical results on the host (C, gcc) and on the
GPU (HIP), so it may not call the platform ``logf``/``sinf``/``cosf``. It uses instead
the polynomials fitted here, evaluated with explicit FMAs only:

* ``ln(1+x)  ~  x + x^2 * P(x)`` for ``x`` in ``[sqrt(1/2)-1, sqrt(2)-1]`` (after the usual
  split ``u = 2^e * m``), degree of ``P`` = 8;
* ``sin(2*pi*r) ~ r * S(r^2)`` and ``cos(2*pi*r) ~ 1 + r^2 * C(r^2)`` for ``r`` in
  ``[-1/8, 1/8]`` (quadrant reduction is done in integers).

Method: weighted least squares on a dense Chebyshev grid (weight = 1/|f| for relative
error), refined by a few Lawson iterations towards the minimax fit, in float64; the
coefficients are then rounded to float32. Running the script prints the coefficients as
C hex-float literals and the max error of the float32 evaluation (in units of float32 ulp
of the true value).

Usage: python rng_poly_fit.py
"""
import numpy as np


def _lawson(A, y, w0, iters=60):
    w = np.ones_like(y)
    c = None
    for _ in range(iters):
        W = np.sqrt(w) * w0
        c, *_ = np.linalg.lstsq(A * W[:, None], y * W, rcond=None)
        err = np.abs((A @ c - y) * w0)
        w = w * (err + 1e-300)
        w /= w.sum()
    return c


def fit_log(deg=8, n=20000):
    lo, hi = np.sqrt(0.5) - 1.0, np.sqrt(2.0) - 1.0
    k = np.arange(n)
    x = (lo + hi) / 2 + (hi - lo) / 2 * np.cos(np.pi * (k + 0.5) / n)
    x = x[np.abs(x) > 1e-6]
    target = (np.log1p(x) - x) / x**2          # P(x)
    A = np.vander(x, deg + 1, increasing=True)
    rel = x**2 / np.abs(np.log1p(x))           # relative weight of P's error in ln(1+x)
    return _lawson(A, target, rel).astype(np.float32)


def fit_sin(n=20000):
    k = np.arange(n)
    r = 0.125 * np.cos(np.pi * (k + 0.5) / n)
    r = r[np.abs(r) > 1e-7]
    target = np.sin(2 * np.pi * r) / r         # S(r^2), terms r^0 .. r^8
    A = np.stack([r**(2 * i) for i in range(5)], axis=1)
    return _lawson(A, target, np.ones_like(r)).astype(np.float32)


def fit_cos(n=20000):
    k = np.arange(n)
    r = 0.125 * np.cos(np.pi * (k + 0.5) / n)
    r = r[np.abs(r) > 1e-7]
    target = (np.cos(2 * np.pi * r) - 1.0) / r**2   # C(r^2), terms r^0 .. r^6
    A = np.stack([r**(2 * i) for i in range(4)], axis=1)
    rel = r**2 / np.cos(2 * np.pi * r)
    return _lawson(A, target, rel).astype(np.float32)


def _fma32(a, b, c):
    return np.float32(np.float64(a) * np.float64(b) + np.float64(c))


def eval_log1p(x, P):
    p = np.float32(P[-1])
    for c in P[-2::-1]:
        p = _fma32(p, x, c)
    x2 = np.float32(np.float64(x) * np.float64(x))
    return _fma32(x2, p, x)


def eval_sin(r, S):
    r2 = np.float32(np.float64(r) * np.float64(r))
    p = np.float32(S[-1])
    for c in S[-2::-1]:
        p = _fma32(p, r2, c)
    return np.float32(np.float64(r) * np.float64(p))


def eval_cos(r, C):
    r2 = np.float32(np.float64(r) * np.float64(r))
    p = np.float32(C[-1])
    for c in C[-2::-1]:
        p = _fma32(p, r2, c)
    return _fma32(r2, p, np.float32(1.0))


def _ulp_err(approx, exact):
    ulp = np.spacing(np.abs(exact).astype(np.float32)).astype(np.float64)
    return np.max(np.abs(approx.astype(np.float64) - exact) / ulp)


def main():
    P, S, C = fit_log(), fit_sin(), fit_cos()
    for name, c in (("LOG1P_P", P), ("SIN_S", S), ("COS_C", C)):
        print(name, ", ".join(float(v).hex() + "f" for v in c))
    m = np.linspace(np.sqrt(0.5), np.sqrt(2.0), 200001).astype(np.float32)
    x = (m - np.float32(1.0)).astype(np.float32)
    print("log1p max err (ulp):", _ulp_err(eval_log1p(x, P), np.log1p(x.astype(np.float64))))
    r = (np.arange(-2**21, 2**21, 7).astype(np.float64) * 2.0**-24).astype(np.float32)
    print("sin max err (ulp):", _ulp_err(eval_sin(r, S), np.sin(2 * np.pi * r.astype(np.float64))))
    print("cos max err (ulp):", _ulp_err(eval_cos(r, C), np.cos(2 * np.pi * r.astype(np.float64))))


if __name__ == "__main__":
    main()
