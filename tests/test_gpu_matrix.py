"""GPU parity of the matrix-core (MFMA) FP32 path kernel.

Per path against the FP64 oracle on the generator's own normals (hjm_generate_normals,
which the path kernels reproduce bit for bit): forward
rates num_x within 2e-6 absolute and discount factors num_y within 1e-6 relative up to
10y, and (like the VALU FP32 kernel) within 5e-5 and 5e-4 at 25y; EE and
CVA of the precision configuration (131072 paths x 10y) within 1e-4 relative of the FP64
kernel; shard additivity; the kernel choice rules (AUTO, VECTOR, MATRIX).
"""
import math

import numpy as np
import pytest

import cva_hjm_cuda_amd as hjm

pytestmark = pytest.mark.gpu


def _flush(msg):
    print(msg, flush=True)


def test_matrix_paths_match_oracle_per_path(oracle):
    """Up to 10y the FP32 curve stays within 2e-6 of FP64 (matrix kernel) and 5e-7 (VALU
    kernel, whose step coefficients (1 - c, c) sum to 1 exactly in FP32). Over 25y the
    scheme's last two tenors form a Jordan block of eigenvalue 1 (the downwind
    extrapolation keeps f_50 - f_49 fixed), so rounding errors grow with the step count
    and advect to every tenor: there both kernels stay within 5e-5 on the forwards and
    5e-4 on the discount factors (measured 1.9e-5 / 1.4e-4 matrix, 1.0e-5 / 7.8e-5 VALU)."""
    m = hjm.benchmark_market()
    for P, expiry, seed in ((777, 10.0, 5), (256, 25.0, 11), (33, 0.5, 2)):
        _flush("[gpu] matrix kernel vs oracle: %d paths x %gy seed %d" % (P, expiry, seed))
        swap = hjm.InterestRateSwap(1e6, 0.025, expiry)
        z = hjm.generate_normals(seed, P, int(round(expiry / 0.01)))
        nx, ny = oracle.reference_paths(m.spot_rates, m.drifts, m.volatilities, expiry, 0.5,
                                        0.01, z)
        D = int(round(expiry / 0.5))
        live = ny != 0
        errs = {}
        for name, pk in (("matrix", hjm.KERNEL_MATRIX), ("vector", hjm.KERNEL_VECTOR)):
            r = hjm.simulate_reference(m, swap, P, seed=seed, precision=hjm.FP32_KAHAN,
                                       keep_paths=True, path_kernel=pk)
            assert np.array_equal(r.num_y != 0, live), name
            assert np.all(r.num_x[:, D + 1:] == 0) and np.all(r.num_x[:, 0] == 0), name
            errs[name] = (np.max(np.abs(r.num_x - nx)),
                          np.max(np.abs(r.num_y[live] / ny[live] - 1)))
        _flush("  max errors: matrix x %.2e y %.2e, vector x %.2e y %.2e"
               % (errs["matrix"] + errs["vector"]))
        if expiry <= 10.0:
            assert errs["matrix"][0] < 2e-6 and errs["matrix"][1] < 1e-6, errs
            assert errs["vector"][0] < 5e-7 and errs["vector"][1] < 1e-6, errs
        else:
            for ex, ey in errs.values():
                assert ex < 5e-5 and ey < 5e-4, errs


def test_matrix_precision_config_vs_fp64():
    _flush("[gpu] matrix kernel at 131072 x 10y vs FP64 and the vector FP32 kernel")
    m = hjm.benchmark_market()
    swap = hjm.InterestRateSwap(1e6, 0.025, 10.0)
    a = hjm.simulate_reference(m, swap, 131072, seed=7)
    b = hjm.simulate_reference(m, swap, 131072, seed=7, precision=hjm.FP32_KAHAN,
                               path_kernel=hjm.KERNEL_MATRIX)
    c = hjm.simulate_reference(m, swap, 131072, seed=7, precision=hjm.FP32_KAHAN,
                               path_kernel=hjm.KERNEL_VECTOR)
    for r in (b, c):
        assert abs(r.cva / a.cva - 1) < 1e-4
        big = a.expected_exposure > 1e-3 * a.expected_exposure.max()
        rel = np.abs(r.expected_exposure[big] / a.expected_exposure[big] - 1)
        assert np.max(rel) < 1e-4
        assert np.all(r.expected_exposure[20:] == 0)
    # sharded runs add up to the whole run
    h1 = hjm.simulate_reference(m, swap, 65536, seed=7, precision=hjm.FP32_KAHAN,
                                path_kernel=hjm.KERNEL_MATRIX)
    h2 = hjm.simulate_reference(m, swap, 65536, seed=7, precision=hjm.FP32_KAHAN,
                                path_offset=65536, path_kernel=hjm.KERNEL_MATRIX)
    s = h1.sums + h2.sums
    nz = b.sums > 0
    assert np.max(np.abs(s[nz] / b.sums[nz] - 1)) < 1e-12
    # AUTO picks the matrix kernel for this configuration
    d = hjm.simulate_reference(m, swap, 131072, seed=7, precision=hjm.FP32_KAHAN)
    assert np.array_equal(d.sums, b.sums)


def test_matrix_kernel_selection_rules(oracle):
    _flush("[gpu] matrix kernel selection rules")
    m = hjm.benchmark_market()
    swap = hjm.InterestRateSwap(1e6, 0.025, 5.0)
    z = oracle.normals(1, 64, 500)
    with pytest.raises(hjm.HJMError):   # host normals
        hjm.simulate_reference(m, swap, 64, normals=z, precision=hjm.FP32_KAHAN,
                               path_kernel=hjm.KERNEL_MATRIX)
    with pytest.raises(hjm.HJMError):   # 25 steps per tenor, not 50
        hjm.simulate_reference(m, swap, 64, dt=0.02, precision=hjm.FP32_KAHAN,
                               path_kernel=hjm.KERNEL_MATRIX)
    r = hjm.simulate_reference(m, swap, 64, dt=0.02, precision=hjm.FP32_KAHAN)
    assert r.cva > 0 and math.isfinite(r.cva)
