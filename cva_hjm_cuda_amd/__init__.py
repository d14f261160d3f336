"""MI355X-native HJM Monte Carlo path generation + swap exposure (EE, PFE, CVA).

The compute runs in libhjm_cva.so (hand-written HIP kernels for gfx950, C ABI in
include/hjm_cva.h); this package is its Python host layer.
"""
from .api import (FP32_KAHAN, FP64, RNG_HOST, RNG_PHILOX, TENORS, HJMError,  # noqa: F401
                  Greeks, InterestRateSwap, Market, ReferencePlan, ReferenceResult, benchmark_market,
                  calculate_exposure_multi_gpu, cva_greeks, cva_weights, device_count, generate_normals,
                  hjm_drift, library, simulate_reference, PortfolioPlan, PortfolioResult,
                  Tuning, simulate_portfolio, autotune_portfolio, portfolio_cva_greeks)
from .api import KERNEL_AUTO, KERNEL_MATRIX, KERNEL_VECTOR, block_map, block_map64  # noqa: F401
from .api import Calibration, box_muller_parts, calibrate_pca, select_kth  # noqa: F401
from .api import NOALLOC_MAX_PATHS, simulate_reference_noalloc  # noqa: F401
from .history import synthetic_history  # noqa: F401
from .portfolio import Portfolio, synthetic_portfolio  # noqa: F401
