"""Pathwise-adjoint CVA Greeks (51 curve + 153 loading + 51 drift sensitivities).

CPU: a tiny FP64 PyTorch restatement of the reference path, differentiated by
torch.autograd, pins the oracle's backward sweep (a third, independent derivative).
GPU: the HIP adjoint (FP32 adjoint state, FP64 accumulation) against the oracle's FP64
adjoint on identical normals; against central finite differences of the GPU's FP64
forward with common random numbers; FP32 forward against FP64 forward; the drift-inclusive
loading sensitivities against finite differences that recompute the Simpson drift; the
torch.autograd.Function wrapper; and the cost of CVA + Greeks against the forward.
"""
import math

import numpy as np
import pytest
import torch

import cva_hjm_cuda_amd as hjm

T = 51


def torch_reference_cva(spot, vols, drift, acc, notional, K, dtau, dt, normals, w):
    """simulation_cpu.h's path and exposure written with differentiable torch ops."""
    P, S, _ = normals.shape
    stride = int(round(dtau / dt))
    D = S // stride
    c, m = dt / dtau, 1.0 / dtau
    V = vols.reshape(3, T)
    f = spot.expand(P, T)
    A = stride * spot.expand(P, T)
    xs, As = {1: f[:, 1]}, {1: A[:, 1]}
    for n in range(1, (D - 1) * stride + 1):
        phi = normals[:, stride + n - 1, :]
        dF = torch.cat([f[:, 1:] - f[:, :-1], (f[:, T - 1] - f[:, T - 2]).unsqueeze(1)], dim=1)
        f = f + drift * dt + math.sqrt(dt) * (phi @ V) + dF * c
        A = A + f
        if n % stride == 0:
            i = n // stride + 1
            xs[i], As[i] = f[:, i], A[:, i]
    cf = [torch.exp(-As[i] * dt) * notional * acc[i] * (m * (torch.exp(xs[i] / m) - 1) - K)
          for i in range(1, D + 1)]
    J = torch.zeros(P, dtype=torch.float64)
    Vt = torch.zeros(P, dtype=torch.float64)
    for t in range(D - 1, -1, -1):
        Vt = Vt + cf[t]
        J = J + w[t] * torch.relu(Vt)
    return J.mean()


def test_torch_autograd_pins_oracle_adjoint(oracle):
    spot = oracle.nelson_siegel()
    vols = oracle.poly_vols()
    drift = oracle.hjm_drift(vols)
    acc = np.full(T, 0.5)
    z = oracle.normals(5, 16, 300)
    w = oracle.cva_weights(T, 0.5, 0.02, 0.4)
    cva, ds, dv, dm = oracle.reference_cva_greeks(spot, drift, vols, acc, 1e6, 0.025, 3.0, 0.5,
                                                  0.01, z, w)
    ts, tv, td = (torch.tensor(a, dtype=torch.float64, requires_grad=True)
                  for a in (spot, vols, drift))
    J = torch_reference_cva(ts, tv, td, torch.tensor(acc), 1e6, 0.025, 0.5, 0.01,
                            torch.tensor(z), torch.tensor(w))
    J.backward()
    assert J.item() == pytest.approx(cva, rel=1e-12)
    for got, ref in ((ts.grad.numpy(), ds), (tv.grad.numpy(), dv), (td.grad.numpy(), dm)):
        assert np.max(np.abs(got - ref)) <= 1e-9 * np.max(np.abs(ref))


# ------------------------------------------------------------------------------ GPU
def _flush(msg):
    print(msg, flush=True)


@pytest.mark.gpu
def test_gpu_adjoint_matches_oracle_adjoint(oracle):
    """512 paths, and 33 and 1 (a partly live workgroup of the seed and sweep kernels)."""
    m = hjm.benchmark_market()
    z = oracle.normals(21, 512, 500)
    w = oracle.cva_weights(T, 0.5, 0.02, 0.4)
    for P in (512, 33, 1):
        _flush("[gpu] adjoint vs oracle adjoint, %d paths 5y FP64 forward, host normals" % P)
        zp = np.ascontiguousarray(z[:P])
        cva, ds, dv, dm = oracle.reference_cva_greeks(m.spot_rates, m.drifts, m.volatilities,
                                                      m.accrual, 1e6, 0.025, 5.0, 0.5, 0.01, zp,
                                                      w)
        g = hjm.cva_greeks(m, hjm.InterestRateSwap(1e6, 0.025, 5.0), P, normals=zp,
                           precision=hjm.FP64)
        assert g.cva == pytest.approx(cva, rel=1e-10)
        for got, ref in ((g.d_spot, ds), (g.d_vol, dv), (g.d_drift, dm)):
            assert np.max(np.abs(got - ref)) <= 2e-5 * np.max(np.abs(ref))


def _cva64(m, swap, paths, seed):
    return hjm.simulate_reference(m, swap, paths, seed=seed, pfe_quantile=0.0).cva


@pytest.mark.gpu
def test_gpu_adjoint_vs_central_fd_common_random_numbers():
    _flush("[gpu] adjoint vs central FD (bump 1e-5, CRN), 4096 paths 10y")
    m = hjm.benchmark_market()
    swap = hjm.InterestRateSwap(1e6, 0.025, 10.0)
    P, seed, h = 4096, 99, 1e-5
    g = hjm.cva_greeks(m, swap, P, seed=seed, precision=hjm.FP64)
    for name, grad in (("spot_rates", g.d_spot), ("volatilities", g.d_vol),
                       ("drifts", g.d_drift)):
        base = getattr(m, name)
        fd = np.zeros_like(grad)
        for j in range(len(base)):
            up = hjm.Market(m.spot_rates, m.drifts, m.volatilities, m.accrual)
            dn = hjm.Market(m.spot_rates, m.drifts, m.volatilities, m.accrual)
            getattr(up, name)[j] += h
            getattr(dn, name)[j] -= h
            fd[j] = (_cva64(up, swap, P, seed) - _cva64(dn, swap, P, seed)) / (2 * h)
        scale = np.max(np.abs(fd))
        err = np.max(np.abs(grad - fd)) / scale
        _flush("  %s: max |adjoint - FD| / max|FD| = %.2e" % (name, err))
        assert err < 2e-3, (name, err)


@pytest.mark.gpu
def test_gpu_total_vol_greeks_include_drift_dependence():
    _flush("[gpu] d_vol_total vs FD with the Simpson drift recomputed")
    m = hjm.benchmark_market()
    swap = hjm.InterestRateSwap(1e6, 0.025, 10.0)
    P, seed, h = 4096, 5, 1e-5
    g = hjm.cva_greeks(m, swap, P, seed=seed, precision=hjm.FP64)
    rng = np.random.default_rng(0)
    for j in rng.choice(3 * T, 24, replace=False):
        vu, vd = m.volatilities.copy(), m.volatilities.copy()
        vu[j] += h
        vd[j] -= h
        up = hjm.Market(m.spot_rates, hjm.hjm_drift(vu), vu, m.accrual)
        dn = hjm.Market(m.spot_rates, hjm.hjm_drift(vd), vd, m.accrual)
        fd = (_cva64(up, swap, P, seed) - _cva64(dn, swap, P, seed)) / (2 * h)
        assert abs(g.d_vol_total[j] - fd) < 2e-3 * np.max(np.abs(g.d_vol_total)), j


@pytest.mark.gpu
def test_gpu_fp32_forward_greeks_close_to_fp64():
    _flush("[gpu] adjoint with FP32 vs FP64 forward, 65536 paths 10y")
    m = hjm.benchmark_market()
    swap = hjm.InterestRateSwap(1e6, 0.025, 10.0)
    a = hjm.cva_greeks(m, swap, 65536, seed=7, precision=hjm.FP64)
    b = hjm.cva_greeks(m, swap, 65536, seed=7, precision=hjm.FP32_KAHAN)
    assert abs(b.cva / a.cva - 1) < 1e-4
    for x, y in ((b.d_spot, a.d_spot), (b.d_vol, a.d_vol), (b.d_drift, a.d_drift)):
        assert np.max(np.abs(x - y)) <= 1e-3 * np.max(np.abs(y))
    # cost against a forward-only plan of the same size (path kernel without stored
    # normals, exposure, reductions, no PFE); the median of a few runs of each
    t = np.array([hjm.cva_greeks(m, swap, 65536, seed=7, precision=hjm.FP32_KAHAN).timings_ms
                  for _ in range(3)])
    plan = hjm.ReferencePlan(m, swap, 65536, precision=hjm.FP32_KAHAN, seed=7, pfe_quantile=0.0)
    plan.execute()
    fwd = float(np.median([plan.execute().timings_ms[4] for _ in range(3)]))
    plan.close()
    ratio = float(np.median(t[:, 2])) / fwd
    _flush("  CVA+Greeks / forward-only device time = %.2f" % ratio)
    # the FP32 forward runs on the matrix cores, the one-pass reverse sweep on the VALU;
    # measured 2.97 on an MI355X; the target is 4
    assert ratio <= 4.0


@pytest.mark.gpu
def test_gpu_fp32_greeks_match_oracle_adjoint(oracle):
    """The default FP32 Greeks (matrix-core forward with in-kernel normals, one-pass sweep
    reading the stored shocks) against the oracle's FP64 adjoint on the same normals
    (hjm_generate_normals gives the in-kernel generator's output): CVA within 1e-5, every
    Greek within 1e-3 of its family's largest."""
    _flush("[gpu] FP32 adjoint vs oracle adjoint, 4096 paths 10y, in-kernel normals")
    m = hjm.benchmark_market()
    P, seed = 4096, 13
    z = hjm.generate_normals(seed, P, 1000)
    w = oracle.cva_weights(T, 0.5, 0.02, 0.4)
    cva, ds, dv, dm = oracle.reference_cva_greeks(m.spot_rates, m.drifts, m.volatilities,
                                                  m.accrual, 1e6, 0.025, 10.0, 0.5, 0.01, z, w)
    g = hjm.cva_greeks(m, hjm.InterestRateSwap(1e6, 0.025, 10.0), P, seed=seed,
                       precision=hjm.FP32_KAHAN)
    assert g.stored_normals_bytes == P * 1000 * 3 * 4
    assert abs(g.cva / cva - 1) < 1e-5
    for got, ref in ((g.d_spot, ds), (g.d_vol, dv), (g.d_drift, dm)):
        assert np.max(np.abs(got - ref)) <= 1e-3 * np.max(np.abs(ref))


@pytest.mark.gpu
def test_gpu_stored_and_regenerated_shocks_give_the_same_greeks():
    """After the matrix-core FP32 forward the sweep reads the forward's stored shocks; asked
    to, it regenerates them from the generator instead: the same Greeks."""
    _flush("[gpu] adjoint: stored vs regenerated shocks, 8192 paths 10y FP32")
    m = hjm.benchmark_market()
    swap = hjm.InterestRateSwap(1e6, 0.025, 10.0)
    a = hjm.cva_greeks(m, swap, 8192, seed=3, precision=hjm.FP32_KAHAN)
    b = hjm.cva_greeks(m, swap, 8192, seed=3, precision=hjm.FP32_KAHAN, regenerate_normals=True)
    assert a.stored_normals_bytes == 8192 * 1000 * 3 * 4 and b.stored_normals_bytes == 0
    assert a.cva == b.cva
    for x, y in ((a.d_spot, b.d_spot), (a.d_vol, b.d_vol), (a.d_drift, b.d_drift)):
        assert np.allclose(x, y, rtol=1e-12, atol=0)


@pytest.mark.gpu
def test_gpu_autograd_wrapper():
    _flush("[gpu] torch.autograd wrapper")
    from cva_hjm_cuda_amd.autograd import cva as cva_fn
    m = hjm.benchmark_market()
    swap = hjm.InterestRateSwap(1e6, 0.025, 5.0)
    s = torch.tensor(m.spot_rates, requires_grad=True)
    v = torch.tensor(m.volatilities, requires_grad=True)
    y = cva_fn(s, v, swap, 8192, seed=3, precision=hjm.FP64)
    (2.0 * y).backward()
    g = hjm.cva_greeks(m, swap, 8192, seed=3, precision=hjm.FP64)
    assert y.item() == pytest.approx(g.cva, rel=1e-14)
    assert np.allclose(s.grad.numpy(), 2.0 * g.d_spot, rtol=1e-14, atol=0)
    assert np.allclose(v.grad.numpy(), 2.0 * g.d_vol_total, rtol=1e-14, atol=0)
    d = torch.tensor(m.drifts, requires_grad=True)
    v2 = torch.tensor(m.volatilities, requires_grad=True)
    y2 = cva_fn(s.detach(), v2, swap, 8192, drift=d, seed=3, precision=hjm.FP64)
    y2.backward()
    assert np.allclose(v2.grad.numpy(), g.d_vol, rtol=1e-14, atol=0)
    assert np.allclose(d.grad.numpy(), g.d_drift, rtol=1e-14, atol=0)
