"""torch.autograd.Function over the C ABI: CVA as a differentiable function of the
initial forward curve, the volatility loadings and (optionally) the drift table.

The forward and the pathwise adjoint both run in libhjm_cva.so on the GPU; torch is
only the interface (the 255 inputs and the scalar output are copied through host
memory). When ``drift`` is omitted it is the cumulative-Simpson HJM drift of ``vols``
and the gradient with respect to ``vols`` includes that dependence.
"""
from __future__ import annotations

import numpy as np
import torch

from . import api


class _CVA(torch.autograd.Function):
    @staticmethod
    def forward(ctx, spot, vols, drift, accrual, swap, paths, dt, precision, seed, hazard,
                recovery, device, derived_drift):
        s = spot.detach().to("cpu", torch.float64).numpy()
        v = vols.detach().to("cpu", torch.float64).numpy().reshape(-1)
        d = (api.hjm_drift(v, swap.dtau) if derived_drift
             else drift.detach().to("cpu", torch.float64).numpy())
        a = accrual.detach().to("cpu", torch.float64).numpy()
        g = api.cva_greeks(api.Market(s, d, v, a), swap, paths, dt, precision, seed,
                           hazard=hazard, recovery=recovery, device=device)
        ctx.derived_drift = derived_drift
        ctx.shapes = (spot.shape, vols.shape)
        ctx.save_for_backward(torch.from_numpy(g.d_spot), torch.from_numpy(
            g.d_vol_total if derived_drift else g.d_vol), torch.from_numpy(g.d_drift))
        ctx.devices = (spot.device, vols.device)
        return torch.tensor(g.cva, dtype=torch.float64, device=spot.device)

    @staticmethod
    def backward(ctx, grad_out):
        ds, dv, dm = ctx.saved_tensors
        go = grad_out.detach().to("cpu", torch.float64)
        gs = (go * ds).reshape(ctx.shapes[0]).to(ctx.devices[0])
        gv = (go * dv).reshape(ctx.shapes[1]).to(ctx.devices[1])
        gd = None if ctx.derived_drift else (go * dm).to(ctx.devices[0])
        return gs, gv, gd, None, None, None, None, None, None, None, None, None, None


def cva(spot: torch.Tensor, vols: torch.Tensor, swap: api.InterestRateSwap, paths: int,
        drift: torch.Tensor | None = None, accrual: torch.Tensor | None = None,
        dt: float = 0.01, precision: int = api.FP32_KAHAN, seed: int = 1234,
        hazard: float = 0.02, recovery: float = 0.4, device: int = 0) -> torch.Tensor:
    """Differentiable CVA (reference semantics, payer IRS) computed on the GPU."""
    if accrual is None:
        accrual = torch.full((api.TENORS,), 0.5, dtype=torch.float64)
    derived = drift is None
    if derived:
        drift = torch.zeros(api.TENORS, dtype=torch.float64)
    return _CVA.apply(spot, vols, drift, accrual, swap, paths, dt, precision, seed, hazard,
                      recovery, device, derived)
