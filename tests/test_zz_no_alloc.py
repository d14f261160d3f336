"""The no-allocation entry point (hjm_simulate_reference_noalloc) through the C ABI.

Its 53 raw sums, written to a caller device buffer on the caller's stream with every
intermediate buffer in the library's static scratch array, equal hjm_simulate_reference's
(same kernels: the one-path-per-lane path kernel, exposure, fixed-order column sums) bit
for bit, on two calls in a row whose second is smaller and leaves the first's data in
the scratch array; sizes past the limit and unsupported options are refused.
"""
import numpy as np
import pytest

import cva_hjm_cuda_amd as hjm

pytestmark = pytest.mark.gpu


def _flush(msg):
    print(msg, flush=True)


def test_noalloc_equals_simulate_reference_on_two_calls():
    import torch
    m = hjm.benchmark_market()
    dev = torch.device("cuda", 0)
    stream = torch.cuda.current_stream(dev).cuda_stream
    cases = [  # (paths, expiry, precision, seed, path_offset): the second call is smaller
        (hjm.NOALLOC_MAX_PATHS, 10.0, hjm.FP64, 11, 0),
        (3001, 5.0, hjm.FP32_KAHAN, 5, 777),
    ]
    for P, expiry, prec, seed, off in cases:
        _flush("[gpu] no-alloc entry point: %d paths x %gy precision %d" % (P, expiry, prec))
        swap = hjm.InterestRateSwap(1e6, 0.025, expiry)
        d_sums = torch.full((53,), float("nan"), dtype=torch.float64, device=dev)
        hjm.simulate_reference_noalloc(m, swap, P, d_sums.data_ptr(), stream, precision=prec,
                                       seed=seed, path_offset=off)
        torch.cuda.synchronize(dev)
        got = d_sums.cpu().numpy()
        want = hjm.simulate_reference(m, swap, P, precision=prec, seed=seed, path_offset=off,
                                      path_kernel=hjm.KERNEL_VECTOR, pfe_quantile=0.0).sums
        assert np.array_equal(got, want), (P, expiry, prec)
        assert got[0] > 0 and np.all(got[int(expiry / 0.5):51] == 0)


def test_noalloc_refuses_what_it_cannot_run():
    import torch
    m = hjm.benchmark_market()
    swap = hjm.InterestRateSwap(1e6, 0.025, 5.0)
    d_sums = torch.zeros(53, dtype=torch.float64, device="cuda")
    for kw, code in ((dict(paths=hjm.NOALLOC_MAX_PATHS + 1), 8),
                     (dict(paths=64, path_kernel=hjm.KERNEL_MATRIX, precision=hjm.FP32_KAHAN), 7)):
        paths = kw.pop("paths")
        with pytest.raises(hjm.HJMError) as ei:
            hjm.simulate_reference_noalloc(m, swap, paths, d_sums.data_ptr(), 0, **kw)
        assert ei.value.code == code
    with pytest.raises(hjm.HJMError) as ei:
        hjm.simulate_reference_noalloc(m, swap, 64, 0, 0)
    assert ei.value.code == 1
