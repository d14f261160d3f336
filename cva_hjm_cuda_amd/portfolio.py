"""Swap portfolios for the batched (portfolio) semantics.

A portfolio is a list of plain-vanilla IRS on the dtau grid: n_s periods (maturity
n_s * dtau), sign +1 payer / -1 receiver, notional, fixed rate, netting set. Netting
sets carry a flat hazard rate each. ``synthetic_portfolio`` draws the benchmark
portfolios: maturities uniform integer years in [2, 30], payer/receiver with p = 0.5,
notionals lognormal(ln 1e6, 1), fixed rates uniform in [1%, 4%], netting sets of uneven
size (multinomial with lognormal weights, every set non-empty), hazards uniform in
[0.5%, 5%], from numpy's PCG64 generator with the given seed.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class Portfolio:
    periods: np.ndarray        # int32 [nswaps]
    sign_notional: np.ndarray  # float64 [nswaps] (+N payer, -N receiver)
    fixed_rate: np.ndarray     # float64 [nswaps]
    set_of: np.ndarray         # int32 [nswaps]
    hazard: np.ndarray         # float64 [nsets]
    dtau: float = 0.5

    @property
    def nswaps(self) -> int:
        return int(self.periods.size)

    @property
    def nsets(self) -> int:
        return int(self.hazard.size)

    def __post_init__(self):
        self.periods = np.ascontiguousarray(self.periods, np.int32)
        self.sign_notional = np.ascontiguousarray(self.sign_notional, np.float64)
        self.fixed_rate = np.ascontiguousarray(self.fixed_rate, np.float64)
        self.set_of = np.ascontiguousarray(self.set_of, np.int32)
        self.hazard = np.ascontiguousarray(self.hazard, np.float64)
        n = self.periods.size
        if not (self.sign_notional.size == self.fixed_rate.size == self.set_of.size == n):
            raise ValueError("swap arrays must have equal length")
        if n and (self.set_of.min() < 0 or self.set_of.max() >= self.hazard.size):
            raise ValueError("netting set index out of range")


def synthetic_portfolio(nswaps: int, nsets: int, seed: int, dtau: float = 0.5) -> Portfolio:
    rng = np.random.default_rng(seed)
    years = rng.integers(2, 31, size=nswaps)
    periods = (years / dtau).round().astype(np.int32)
    sign = np.where(rng.random(nswaps) < 0.5, 1.0, -1.0)
    notional = np.exp(np.log(1e6) + rng.normal(0.0, 1.0, size=nswaps))
    fixed = rng.uniform(0.01, 0.04, size=nswaps)
    weights = np.exp(rng.normal(0.0, 1.0, size=nsets))
    set_of = np.empty(nswaps, np.int32)
    set_of[:nsets] = rng.permutation(nsets)[: min(nsets, nswaps)]
    if nswaps > nsets:
        set_of[nsets:] = rng.choice(nsets, size=nswaps - nsets, p=weights / weights.sum())
    hazard = rng.uniform(0.005, 0.05, size=nsets)
    return Portfolio(periods, sign * notional, fixed, set_of, hazard, dtau)
