"""GPU parity of the FP64 matrix-core path kernel (v_mfma_f64_16x16x4f64 over the 50-step
block map; ``precision=FP64, path_kernel=KERNEL_MATRIX``).

The kernel composes 50 Euler steps into one FP64 affine map, so its rounding differs from
the step-by-step oracle only in order: per path against the FP64 oracle on the
generator's own normals (which the kernel reproduces bit for bit), forward rates within
1e-13 absolute and discount factors within 1e-13 relative up to 15y (measured 4e-15 and
2e-14), 1e-11 at 25y (measured 1.9e-13 and 1.3e-12, the vector FP64 kernel 2.6e-13 on the
forwards: the scheme's Jordan block, see test_gpu_matrix.py, grows every rounding
difference with the step count); EE, PFE and CVA of the precision configuration (131072
paths x 10y) within 1e-12 relative of the vector FP64 kernel (measured 1.6e-14); shard
additivity; the Greeks from its stored shocks against the vector FP64 plan's regenerated
ones.
"""
import numpy as np
import pytest

import cva_hjm_cuda_amd as hjm

pytestmark = pytest.mark.gpu


def _flush(msg):
    print(msg, flush=True)


def test_matrix64_paths_match_oracle_per_path(oracle):
    m = hjm.benchmark_market()
    # 0.5y: no evolved block (4 M-tiles); 10y: 5 tiles, all in LDS; 15y: 6 and 25y: 7
    # tiles, the last ones read from global memory
    for P, expiry, seed in ((777, 10.0, 5), (300, 15.0, 3), (256, 25.0, 11), (33, 0.5, 2)):
        _flush("[gpu] FP64 matrix kernel vs oracle: %d paths x %gy seed %d" % (P, expiry, seed))
        swap = hjm.InterestRateSwap(1e6, 0.025, expiry)
        z = hjm.generate_normals(seed, P, int(round(expiry / 0.01)))
        nx, ny = oracle.reference_paths(m.spot_rates, m.drifts, m.volatilities, expiry, 0.5,
                                        0.01, z)
        D = int(round(expiry / 0.5))
        live = ny != 0
        r = hjm.simulate_reference(m, swap, P, seed=seed, precision=hjm.FP64, keep_paths=True,
                                   path_kernel=hjm.KERNEL_MATRIX)
        v = hjm.simulate_reference(m, swap, P, seed=seed, precision=hjm.FP64, keep_paths=True,
                                   path_kernel=hjm.KERNEL_VECTOR)
        assert np.array_equal(r.num_y != 0, live)
        assert np.all(r.num_x[:, D + 1:] == 0) and np.all(r.num_x[:, 0] == 0)
        ex = np.max(np.abs(r.num_x - nx))
        ey = np.max(np.abs(r.num_y[live] / ny[live] - 1))
        vx = np.max(np.abs(v.num_x - nx))
        _flush("  max errors: x %.2e y %.2e (vector kernel x %.2e)" % (ex, ey, vx))
        tol = 1e-13 if expiry <= 15.0 else 1e-11
        assert ex < tol and ey < tol, (ex, ey)
        assert abs(r.cva - v.cva) <= 1e-11 * abs(v.cva)


def test_matrix64_precision_config_vs_vector_fp64():
    _flush("[gpu] FP64 matrix kernel at 131072 x 10y vs the vector FP64 kernel")
    m = hjm.benchmark_market()
    swap = hjm.InterestRateSwap(1e6, 0.025, 10.0)
    a = hjm.simulate_reference(m, swap, 131072, seed=7, path_kernel=hjm.KERNEL_VECTOR)
    b = hjm.simulate_reference(m, swap, 131072, seed=7, path_kernel=hjm.KERNEL_MATRIX)
    big = a.expected_exposure > 1e-3 * a.expected_exposure.max()
    ee = np.max(np.abs(b.expected_exposure[big] / a.expected_exposure[big] - 1))
    pk = a.pfe > 0
    pf = np.max(np.abs(b.pfe[pk] / a.pfe[pk] - 1))
    _flush("  EE %.2e PFE %.2e CVA %.2e" % (ee, pf, abs(b.cva / a.cva - 1)))
    assert abs(b.cva / a.cva - 1) < 1e-12
    assert ee < 1e-12 and pf < 1e-12
    assert np.all(b.expected_exposure[20:] == 0)
    # sharded runs add up to the whole run
    h1 = hjm.simulate_reference(m, swap, 65536, seed=7, path_kernel=hjm.KERNEL_MATRIX)
    h2 = hjm.simulate_reference(m, swap, 65536, seed=7, path_offset=65536,
                                path_kernel=hjm.KERNEL_MATRIX)
    s = h1.sums + h2.sums
    nz = b.sums > 0
    assert np.max(np.abs(s[nz] / b.sums[nz] - 1)) < 1e-12
    # a ragged path count (last workgroup partly live) equals the same paths of a larger run
    c = hjm.simulate_reference(m, swap, 1000, seed=7, keep_paths=True,
                               path_kernel=hjm.KERNEL_MATRIX)
    d = hjm.simulate_reference(m, swap, 1024, seed=7, keep_paths=True,
                               path_kernel=hjm.KERNEL_MATRIX)
    assert np.array_equal(c.num_x, d.num_x[:1000]) and np.array_equal(c.num_y, d.num_y[:1000])


def test_matrix64_greeks_from_stored_shocks():
    """The forward variants that keep their shocks (4, 5 and 7 M-tiles); the reverse sweep
    is FP32, so the forwards' 1e-14 differences mostly vanish in its rounding."""
    m = hjm.benchmark_market()
    for expiry, P in ((10.0, 16384), (5.0, 4096), (25.0, 4096)):
        _flush("[gpu] FP64 matrix-core Greeks vs the vector FP64 plan, %d paths x %gy"
               % (P, expiry))
        swap = hjm.InterestRateSwap(1e6, 0.025, expiry)
        a = hjm.cva_greeks(m, swap, P, seed=3, precision=hjm.FP64, path_kernel=hjm.KERNEL_VECTOR)
        b = hjm.cva_greeks(m, swap, P, seed=3, precision=hjm.FP64, path_kernel=hjm.KERNEL_MATRIX)
        assert a.stored_normals_bytes == 0
        assert b.stored_normals_bytes == int(round(expiry / 0.01)) * 3 * P * 4
        assert abs(b.cva / a.cva - 1) < 1e-11
        for x, y in ((a.d_spot, b.d_spot), (a.d_vol, b.d_vol), (a.d_drift, b.d_drift)):
            err = np.max(np.abs(x - y)) / np.max(np.abs(x))
            _flush("  max rel difference %.2e" % err)
            assert err < 1e-6


def test_matrix64_selection_rules(oracle):
    _flush("[gpu] FP64 matrix kernel selection rules")
    m = hjm.benchmark_market()
    swap = hjm.InterestRateSwap(1e6, 0.025, 5.0)
    z = oracle.normals(1, 64, 500)
    with pytest.raises(hjm.HJMError):   # host normals
        hjm.simulate_reference(m, swap, 64, normals=z, path_kernel=hjm.KERNEL_MATRIX)
    with pytest.raises(hjm.HJMError):   # 25 steps per tenor, not 50
        hjm.simulate_reference(m, swap, 64, dt=0.02, path_kernel=hjm.KERNEL_MATRIX)
    # AUTO keeps FP64 on the vector kernel (the reference's step order)
    a = hjm.simulate_reference(m, swap, 64, seed=4)
    v = hjm.simulate_reference(m, swap, 64, seed=4, path_kernel=hjm.KERNEL_VECTOR)
    assert np.array_equal(a.sums, v.sums)
