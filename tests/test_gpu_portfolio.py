"""GPU parity of the portfolio semantics with the oracle, and its execution modes.

FP64 on identical normals (the generator's own, hjm_generate_normals): per-path
netting-set exposures, EE, CVA and SE within 1e-10 relative of the oracle; PFE the exact
order statistic; in-kernel Philox bit-identical to the same normals supplied from the host; hipGraph replays bit-identical to eager runs (several replays with eager
work in between); tuned persistent grids and every launch shape of the matrix-core
kernels bit-identical to the defaults, the autotuner's choice replayed bit for bit; FP32 versus
FP64 at the batched configuration's size and for the largest book the ABI takes (4096 swaps
in 512 sets), forward and Greeks; shard additivity.
"""
import dataclasses
import math

import numpy as np
import pytest

import cva_hjm_cuda_amd as hjm
from cva_hjm_cuda_amd.portfolio import Portfolio

pytestmark = pytest.mark.gpu


def _flush(msg):
    print(msg, flush=True)


def _stats_from_exposures(ex, pf, dtau, recovery, oracle, q=0.975):
    P, G, D = ex.shape
    ee = ex.mean(axis=0)
    k = max(1, min(P, math.ceil(q * P)))
    pfe = np.sort(ex, axis=0)[k - 1]
    cva, se = np.zeros(G), np.zeros(G)
    for g in range(G):
        w = oracle.cva_weights(D, dtau, pf.hazard[g], recovery)
        per = ex[:, g, :] @ w
        cva[g] = ee[g] @ w
        se[g] = per.std(ddof=1) / math.sqrt(P)
    return ee, pfe, cva, se


@pytest.fixture(scope="module")
def small(oracle):
    _flush("[gpu] portfolio oracle run: 128 paths x 30y, 64 swaps, 8 sets")
    m = hjm.benchmark_market()
    pf = hjm.synthetic_portfolio(64, 8, seed=7)
    z = hjm.generate_normals(31, 128, 3000)
    ex = oracle.portfolio_exposure(m.spot_rates, m.drifts, m.volatilities, 30.0, 0.5, 0.01, z, pf)
    return m, pf, z, ex


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def test_portfolio_fp64_matches_oracle(small, oracle):
    m, pf, z, ex_o = small
    r = hjm.simulate_portfolio(m, pf, 128, horizon=30.0, precision=hjm.FP64, normals=z,
                               keep_paths=True)
    assert _rel(r.exposures, ex_o) < 1e-10
    ee, pfe, cva, se = _stats_from_exposures(ex_o, pf, 0.5, 0.4, oracle)
    assert _rel(r.expected_exposure, ee) < 1e-10
    assert np.max(np.abs(r.cva / cva - 1)) < 1e-10
    assert np.max(np.abs(r.cva_stderr / se - 1)) < 1e-8
    assert _rel(r.pfe, pfe) < 1e-10
    # the exact order statistic of the GPU's own exposures
    k = math.ceil(0.975 * 128)
    assert np.array_equal(r.pfe, np.sort(r.exposures, axis=0)[k - 1])
    # t = 0 is deterministic
    assert np.all(r.exposures[:, :, 0] == r.exposures[0:1, :, 0])


def test_portfolio_tiny_path_counts(small, oracle):
    """1 and 3 paths: FP64 on host normals equals the oracle per path and in EE, CVA and its
    standard error (0 for one path); PFE is the exact order statistic; the FP32 matrix-core
    passes give those paths the exposures of the first paths of a 64-path run, bit for
    bit."""
    m, pf, z, ex_o = small
    scale = np.max(np.abs(ex_o))
    for P in (1, 3):
        _flush("[gpu] portfolio, %d paths: FP64 host normals vs oracle" % P)
        r = hjm.simulate_portfolio(m, pf, P, horizon=30.0, precision=hjm.FP64, normals=z[:P],
                                   keep_paths=True)
        assert np.max(np.abs(r.exposures - ex_o[:P])) < 1e-10 * scale
        ee = ex_o[:P].mean(axis=0)
        assert np.max(np.abs(r.expected_exposure - ee)) < 1e-10 * scale
        k = math.ceil(0.975 * P)
        assert np.array_equal(r.pfe, np.sort(r.exposures, axis=0)[k - 1])
        for g in range(pf.nsets):
            w = oracle.cva_weights(61, 0.5, pf.hazard[g], 0.4)
            per = ex_o[:P, g, :] @ w
            assert abs(r.cva[g] - per.mean()) <= 1e-10 * max(abs(per.mean()), 1e-300)
            if P == 1:
                assert r.cva_stderr[g] == 0.0
            else:
                se = per.std(ddof=1) / math.sqrt(P)
                # (sum of squares - square of sum / P): rounding at the scale of the mean
                assert abs(r.cva_stderr[g] - se) <= 1e-8 * se + 1e-10 * abs(per.mean())
    _flush("[gpu] portfolio, 1 and 3 paths, FP32 vs the first paths of 64")
    big = hjm.simulate_portfolio(m, pf, 64, horizon=30.0, seed=9, keep_paths=True)
    for P in (1, 3):
        s = hjm.simulate_portfolio(m, pf, P, horizon=30.0, seed=9, keep_paths=True)
        assert np.array_equal(s.exposures, big.exposures[:P])


def test_portfolio_short_horizons(oracle):
    """Horizons of 0.5y (two dates, one evolved block), 1y and 2.5y, swaps cut to the
    horizon: FP64 on the generator's normals equals the oracle per path (1e-10 of scale);
    the FP32 matrix-core passes on the same in-kernel normals within 2e-6 of each set's
    gross notional per path; the EE rows after the last payment are zero."""
    m = hjm.benchmark_market()
    base = hjm.synthetic_portfolio(24, 5, seed=9)
    for horizon in (0.5, 1.0, 2.5):
        _flush("[gpu] portfolio horizon %gy: FP64 vs oracle, FP32 vs FP64" % horizon)
        E = int(round(horizon / 0.5))
        pf = Portfolio(np.minimum(base.periods, E), base.sign_notional, base.fixed_rate,
                       base.set_of, base.hazard)
        z = hjm.generate_normals(13, 40, int(round(horizon / 0.01)))
        ex_o = oracle.portfolio_exposure(m.spot_rates, m.drifts, m.volatilities, horizon, 0.5,
                                         0.01, z, pf)
        a = hjm.simulate_portfolio(m, pf, 40, horizon=horizon, precision=hjm.FP64, normals=z,
                                   keep_paths=True)
        assert _rel(a.exposures, ex_o) < 1e-10
        b = hjm.simulate_portfolio(m, pf, 40, horizon=horizon, seed=13, keep_paths=True)
        gross = np.bincount(pf.set_of, weights=np.abs(pf.sign_notional), minlength=pf.nsets)
        assert np.all(np.abs(b.exposures - a.exposures) <= 2e-6 * gross[None, :, None])
        assert np.all(a.expected_exposure[:, -1] == 0) and np.all(b.expected_exposure[:, -1] == 0)


def test_portfolio_fp64_matrix_path_kernel(oracle):
    """The FP64 path kernel on the matrix cores (the portfolio block map in FP64,
    path_kernel=MATRIX): exposures within 1e-12 of the FP64 oracle's scale on the
    generator's own normals (200 paths x 30y); EE, PFE and CVA of 20000 paths (a ragged
    last workgroup) within 1e-12 relative of the vector FP64 kernel; the Greeks from its
    stored shocks against the vector plan's regenerated ones; AUTO keeps FP64 on the vector
    kernel; the autotuner leaves the matrix-core plan at its defaults."""
    _flush("[gpu] portfolio FP64 matrix-core path kernel vs oracle, 200 paths x 30y")
    m = hjm.benchmark_market()
    pf = hjm.synthetic_portfolio(48, 12, seed=6)
    z = hjm.generate_normals(21, 200, 3000)
    ex_o = oracle.portfolio_exposure(m.spot_rates, m.drifts, m.volatilities, 30.0, 0.5, 0.01,
                                     z, pf)
    r = hjm.simulate_portfolio(m, pf, 200, horizon=30.0, precision=hjm.FP64, seed=21,
                               keep_paths=True, path_kernel=hjm.KERNEL_MATRIX)
    err = _rel(r.exposures, ex_o)
    _flush("  exposures: max error %.2e of scale" % err)
    assert err < 1e-12
    _flush("[gpu] portfolio FP64 matrix-core vs vector path kernel, 20000 paths")
    a = hjm.simulate_portfolio(m, pf, 20000, horizon=30.0, precision=hjm.FP64, seed=5,
                               path_kernel=hjm.KERNEL_VECTOR)
    b = hjm.simulate_portfolio(m, pf, 20000, horizon=30.0, precision=hjm.FP64, seed=5,
                               path_kernel=hjm.KERNEL_MATRIX)
    ee, pk = a.expected_exposure > 0, a.pfe > 0
    d = (np.max(np.abs(b.expected_exposure[ee] / a.expected_exposure[ee] - 1)),
         np.max(np.abs(b.pfe[pk] / a.pfe[pk] - 1)), np.max(np.abs(b.cva / a.cva - 1)))
    _flush("  EE %.2e PFE %.2e CVA %.2e" % d)
    assert max(d) < 1e-12
    c = hjm.simulate_portfolio(m, pf, 20000, horizon=30.0, precision=hjm.FP64, seed=5)
    assert np.array_equal(c.expected_exposure, a.expected_exposure)
    _flush("[gpu] portfolio FP64 Greeks, matrix-core forward with stored shocks")
    ga = hjm.portfolio_cva_greeks(m, pf, 4096, horizon=30.0, precision=hjm.FP64, seed=8,
                                  path_kernel=hjm.KERNEL_VECTOR)
    gb = hjm.portfolio_cva_greeks(m, pf, 4096, horizon=30.0, precision=hjm.FP64, seed=8,
                                  path_kernel=hjm.KERNEL_MATRIX)
    assert ga.stored_normals_bytes == 0 and gb.stored_normals_bytes == 4096 * 3000 * 3 * 4
    assert abs(gb.cva / ga.cva - 1) < 1e-11
    for x, y in ((gb.d_spot, ga.d_spot), (gb.d_vol, ga.d_vol), (gb.d_drift, ga.d_drift)):
        assert np.max(np.abs(x - y)) <= 1e-6 * np.max(np.abs(y))
    # nothing to tune on the FP64 matrix-core path kernel: the defaults, with a graph
    t, ms = hjm.autotune_portfolio(m, pf, 4096, horizon=30.0, precision=hjm.FP64,
                                   path_kernel=hjm.KERNEL_MATRIX)
    assert (t.block_size, t.blocks_per_cu, t.use_graph) == (256, 0, 1) and ms > 0


def test_portfolio_fp32_pfe_exact_on_long_rows():
    """PFE over 132072 FP32 / 66000 FP64 paths (the sampled-bracket selection, counted in
    the statistics pass): the exact order statistic of the plan's own exposures, for every
    set and date."""
    _flush("[gpu] portfolio FP32 PFE, 132072 paths x 30y, 16 swaps, 4 sets")
    m = hjm.benchmark_market()
    pf = hjm.synthetic_portfolio(16, 4, seed=3)
    P = 132072  # the last statistics block is partial
    r = hjm.simulate_portfolio(m, pf, P, horizon=30.0, precision=hjm.FP32_KAHAN, seed=11,
                               keep_paths=True)
    k = math.ceil(0.975 * P)
    want = np.sort(r.exposures.astype(np.float32), axis=0)[k - 1].astype(np.float64)
    assert np.array_equal(r.pfe, want)
    assert np.all(r.pfe[:, 1] > 0)
    # the median: a bracket too wide to stage all 61 dates in the statistics pass's LDS,
    # so the staging slices are flushed every 36 dates
    r = hjm.simulate_portfolio(m, pf, P, horizon=30.0, precision=hjm.FP32_KAHAN, seed=11,
                               keep_paths=True, pfe_quantile=0.5)
    k = math.ceil(0.5 * P)
    want = np.sort(r.exposures.astype(np.float32), axis=0)[k - 1].astype(np.float64)
    assert np.array_equal(r.pfe, want)
    # FP64 rows (sample of 8192 paths, a partial last statistics block), counted in the
    # statistics pass: 10y horizon (21 dates, staged whole) and 30y at the 90% quantile
    # (staging slices flushed every 17 dates)
    P = 66000
    pf10 = Portfolio(np.minimum(pf.periods, 20), pf.sign_notional, pf.fixed_rate, pf.set_of,
                     pf.hazard)
    for book, horizon, q in ((pf10, 10.0, 0.975), (pf, 30.0, 0.9)):
        r = hjm.simulate_portfolio(m, book, P, horizon=horizon, precision=hjm.FP64, seed=11,
                                   keep_paths=True, pfe_quantile=q)
        k = math.ceil(q * P)
        assert np.array_equal(r.pfe, np.sort(r.exposures, axis=0)[k - 1]), (horizon, q)


def test_fused_statistics_equal_the_exposure_pass():
    """FP32 plans can reduce the exposures inside the valuation (EE, per-path CVA moments,
    the PFE bracket counting; tuning.fuse_statistics) unless the per-path exposures are
    asked for. On one plan, 160 sets (two set groups of the fused pass, 128 + 32) x 40000
    paths (a ragged last workgroup) x 30y: the fused run's PFE equals the exposure pass's
    (the exact order statistic of the same exposures), EE and CVA to 1e-6, the standard
    error to FP32 accumulation of the per-path CVA; and with one candidate slot per set
    and date the rows overflow, are valued again whole, and the PFE is still exact. At the
    97.5 % and 50 % quantiles. (EE and CVA: FP32 partial sums of 32 paths in the fused
    pass, FP64 per path in the exposure pass.)"""
    m = hjm.benchmark_market()
    pf = hjm.synthetic_portfolio(320, 160, seed=5)
    P = 40000
    for q in (0.975, 0.5):
        _flush("[gpu] fused statistics vs exposure pass, q = %g" % q)
        plan = hjm.PortfolioPlan(m, pf, P, horizon=30.0, seed=9, pfe_quantile=q,
                                 tuning=hjm.Tuning(fuse_statistics=True))
        a = plan.execute(keep_paths=True)
        b = plan.execute()
        plan.close()
        k = math.ceil(q * P)
        want = np.sort(a.exposures.astype(np.float32), axis=0)[k - 1].astype(np.float64)
        assert np.array_equal(a.pfe, want)
        assert np.array_equal(b.pfe, a.pfe)
        assert np.allclose(b.expected_exposure, a.expected_exposure, rtol=1e-6, atol=0)
        assert np.allclose(b.cva, a.cva, rtol=1e-6, atol=0)
        assert np.allclose(b.cva_stderr, a.cva_stderr, rtol=1e-5, atol=0)
        plan = hjm.PortfolioPlan(m, pf, P, horizon=30.0, seed=9, pfe_quantile=q,
                                 tuning=hjm.Tuning(fuse_statistics=True, pfe_stage_slots=1))
        c = plan.execute()
        plan.close()
        assert np.array_equal(c.pfe, a.pfe)
        assert np.array_equal(c.expected_exposure, b.expected_exposure)


def test_portfolio_in_kernel_rng_equals_host_normals(small):
    m, pf, z, _ = small
    a = hjm.simulate_portfolio(m, pf, 128, horizon=30.0, precision=hjm.FP64, normals=z)
    b = hjm.simulate_portfolio(m, pf, 128, horizon=30.0, precision=hjm.FP64, seed=31)
    assert np.array_equal(a.expected_exposure, b.expected_exposure)
    assert np.array_equal(a.cva, b.cva) and np.array_equal(a.pfe, b.pfe)


def test_portfolio_matrix_kernel_matches_oracle(small):
    _flush("[gpu] portfolio matrix-core FP32 path kernel vs FP64 oracle, 128 paths x 30y")
    m, pf, z, ex_o = small
    r = hjm.simulate_portfolio(m, pf, 128, horizon=30.0, precision=hjm.FP32_KAHAN, seed=31,
                               keep_paths=True, path_kernel=hjm.KERNEL_MATRIX)
    v = hjm.simulate_portfolio(m, pf, 128, horizon=30.0, precision=hjm.FP32_KAHAN, seed=31,
                               keep_paths=True, path_kernel=hjm.KERNEL_VECTOR)
    scale = np.max(np.abs(ex_o))
    assert np.max(np.abs(r.exposures - ex_o)) <= 1e-4 * scale
    assert np.max(np.abs(v.exposures - ex_o)) <= 1e-4 * scale
    assert np.all(r.exposures[:, :, 0] == r.exposures[0:1, :, 0])
    with pytest.raises(hjm.HJMError):   # the matrix cores draw their own normals
        hjm.simulate_portfolio(m, pf, 128, horizon=30.0, normals=z,
                               path_kernel=hjm.KERNEL_MATRIX)


def test_graph_replays_and_tuned_grids_are_bit_identical():
    _flush("[gpu] hipGraph replays vs eager, tuned grids (70000 paths: bracket PFE in the graph)")
    m = hjm.benchmark_market()
    pf = hjm.synthetic_portfolio(96, 12, seed=3)
    eager = hjm.simulate_portfolio(m, pf, 70000, horizon=30.0, seed=9)
    plan = hjm.PortfolioPlan(m, pf, 70000, horizon=30.0, seed=9,
                             tuning=hjm.Tuning(256, 0, use_graph=True))
    for rep in range(4):
        r = plan.execute()
        assert np.array_equal(r.expected_exposure, eager.expected_exposure), rep
        assert np.array_equal(r.pfe, eager.pfe) and np.array_equal(r.cva, eager.cva), rep
        # unrelated eager work on the device between replays
        hjm.simulate_reference(m, hjm.InterestRateSwap(1e6, 0.02, 5.0), 512, seed=rep)
    plan.close()
    # persistent grids of the VALU path kernel
    eager_v = hjm.simulate_portfolio(m, pf, 70000, horizon=30.0, seed=9,
                                     path_kernel=hjm.KERNEL_VECTOR)
    for t in (hjm.Tuning(64, 1), hjm.Tuning(128, 4), hjm.Tuning(256, 2, use_graph=True)):
        r = hjm.simulate_portfolio(m, pf, 70000, horizon=30.0, seed=9, tuning=t,
                                   path_kernel=hjm.KERNEL_VECTOR)
        assert np.array_equal(r.expected_exposure, eager_v.expected_exposure), t
    assert np.max(np.abs(eager_v.cva / eager.cva - 1)) < 1e-4
    tune, ms = hjm.autotune_portfolio(m, pf, 70000, horizon=30.0, seed=9,
                                      path_kernel=hjm.KERNEL_VECTOR)
    assert tune.block_size in (64, 128, 256) and tune.blocks_per_cu in (0, 1, 2, 4, 8)
    assert tune.use_graph and ms > 0


def _matrix_shapes():
    """Every launch shape of the matrix-core path kernel (x persistent grid) and of the
    valuation (x LDS stage) that hjm_tuning accepts."""
    for tiles, waves in ((2, 8), (2, 12), (1, 12), (1, 16)):
        for bpc in (0, 1):
            yield hjm.Tuning(blocks_per_cu=bpc, path_tiles=tiles, path_waves=waves)
    for tiles in (2, 4):
        for waves in (4, 8):
            for stage in (0, 1, 3):
                yield hjm.Tuning(value_tiles=tiles, value_waves=waves, value_stage=stage)


def test_matrix_shapes_autotune_and_tuned_graph_are_bit_identical():
    """The matrix-core kernels' launch shapes change no result bit (each path and each
    exposure is computed by the same instructions whatever the tiling), the autotuner
    returns one of them, and the hipGraph captured at the tuned shape replays bit for bit
    what an eager run at that shape gives. 40 sets (3 M-tiles: a stage of 1 or 3 splits
    them), 70000 paths (not a multiple of any group: ragged last groups)."""
    _flush("[gpu] matrix-core launch shapes, autotuner, graph at the tuned shape")
    m = hjm.benchmark_market()
    pf = hjm.synthetic_portfolio(120, 40, seed=5)
    P = 70000
    base = hjm.simulate_portfolio(m, pf, P, horizon=30.0, seed=9)
    for t in _matrix_shapes():
        r = hjm.simulate_portfolio(m, pf, P, horizon=30.0, seed=9, tuning=t)
        assert np.array_equal(r.expected_exposure, base.expected_exposure), t
        assert np.array_equal(r.pfe, base.pfe) and np.array_equal(r.cva, base.cva), t
        assert np.array_equal(r.cva_stderr, base.cva_stderr), t
    with pytest.raises(hjm.HJMError):
        hjm.simulate_portfolio(m, pf, 1024, horizon=30.0, seed=9,
                               tuning=hjm.Tuning(path_tiles=1, path_waves=8))
    with pytest.raises(hjm.HJMError):
        hjm.simulate_portfolio(m, pf, 1024, horizon=30.0, seed=9,
                               tuning=hjm.Tuning(value_tiles=3))
    tune, ms = hjm.autotune_portfolio(m, pf, P, horizon=30.0, seed=9)
    assert tune.use_graph and ms > 0
    assert (tune.path_tiles, tune.path_waves) in ((2, 8), (2, 12), (1, 12), (1, 16)), tune
    assert tune.blocks_per_cu in (0, 1) and tune.value_tiles in (2, 4), tune
    assert tune.value_waves in (4, 8) and tune.value_stage in (2, 4, 8), tune
    eager_t = dataclasses.replace(tune, use_graph=False)
    ref = hjm.simulate_portfolio(m, pf, P, horizon=30.0, seed=9, tuning=eager_t)
    if not tune.fuse_statistics:
        assert np.array_equal(ref.expected_exposure, base.expected_exposure)
    plan = hjm.PortfolioPlan(m, pf, P, horizon=30.0, seed=9, tuning=tune)
    for rep in range(3):
        r = plan.execute()
        assert np.array_equal(r.expected_exposure, ref.expected_exposure), rep
        assert np.array_equal(r.pfe, ref.pfe) and np.array_equal(r.cva, ref.cva), rep
        hjm.simulate_reference(m, hjm.InterestRateSwap(1e6, 0.02, 5.0), 512, seed=rep)
    plan.close()


def test_batched_config_fp32_vs_fp64_and_shards():
    _flush("[gpu] batched config: 262144 paths x 30y, 256 swaps, 32 sets, FP32 vs FP64")
    m = hjm.benchmark_market()
    pf = hjm.synthetic_portfolio(256, 32, seed=2024)
    a = hjm.simulate_portfolio(m, pf, 262144, horizon=30.0, precision=hjm.FP64, seed=2024)
    b = hjm.simulate_portfolio(m, pf, 262144, horizon=30.0, precision=hjm.FP32_KAHAN, seed=2024)
    assert np.all(a.cva > 0) and np.all(a.cva_stderr < 0.1 * a.cva)
    assert abs(b.cva.sum() / a.cva.sum() - 1) < 1e-4
    assert np.max(np.abs(b.cva / a.cva - 1)) < 1e-3
    assert np.all(a.pfe >= 0) and np.all(a.expected_exposure[:, -1] == 0)
    h1 = hjm.simulate_portfolio(m, pf, 131072, horizon=30.0, precision=hjm.FP64, seed=2024)
    h2 = hjm.simulate_portfolio(m, pf, 131072, horizon=30.0, precision=hjm.FP64, seed=2024,
                                path_offset=131072)
    ee = 0.5 * (h1.expected_exposure + h2.expected_exposure)
    assert _rel(ee, a.expected_exposure) < 1e-12


def test_largest_book_fp32_against_fp64_forward_and_greeks():
    """The largest book the ABI takes, 4096 swaps in 512 netting sets (32 set tiles: the
    valuation stages them 8 at a time, the matrix-core adjoint seeds take them in four
    groups), and 600 swaps in 200 sets (13 tiles: a partial second seed group); 16384
    paths x 30y. FP32 against FP64 on the same in-kernel normals: total CVA within 1e-5, every
    set's CVA within 1e-4, EE within 2e-6 of the set's gross notional (FP32 rounds a set's
    value at the scale of its swaps' notionals, which netting can leave far above its
    exposure: measured 2.6e-7 of the gross, up to 1.1e-5 of a set's peak EE; CVA 6.9e-6);
    CVA + 204 Greeks of the book: the CVA equal to the forward's, the Greeks within 1e-4 x
    scale (measured 7.8e-6); at 512 sets the fused valuation-statistics pass (four set groups)
    equal to the exposure pass: PFE exactly, EE and CVA to 1e-6."""
    m = hjm.benchmark_market()
    P = 16384
    for nsw, nsets in ((4096, 512), (600, 200)):
        _flush("[gpu] %d swaps in %d sets, %d paths x 30y: FP32 vs FP64, forward and Greeks"
               % (nsw, nsets, P))
        pf = hjm.synthetic_portfolio(nsw, nsets, seed=77)
        a = hjm.simulate_portfolio(m, pf, P, horizon=30.0, precision=hjm.FP64, seed=5)
        b = hjm.simulate_portfolio(m, pf, P, horizon=30.0, precision=hjm.FP32_KAHAN, seed=5)
        assert np.all(np.isfinite(b.cva)) and np.all(np.isfinite(b.expected_exposure))
        assert np.all(a.cva >= 0) and a.cva.max() > 0
        assert abs(b.cva.sum() / a.cva.sum() - 1) < 1e-5
        assert np.all(np.abs(b.cva - a.cva) <= 1e-4 * a.cva)
        gross = np.bincount(pf.set_of, weights=np.abs(pf.sign_notional), minlength=nsets)
        assert np.all(np.abs(b.expected_exposure - a.expected_exposure) <= 2e-6 * gross[:, None])
        if nsets == 512:
            # the fused valuation-statistics pass over four groups of 128 sets
            plan = hjm.PortfolioPlan(m, pf, P, horizon=30.0, seed=5,
                                     tuning=hjm.Tuning(fuse_statistics=True))
            c = plan.execute()
            plan.close()
            assert np.array_equal(c.pfe, b.pfe)
            assert np.allclose(c.expected_exposure, b.expected_exposure, rtol=1e-6, atol=0)
            assert np.allclose(c.cva, b.cva, rtol=1e-6, atol=0)
        g32 = hjm.portfolio_cva_greeks(m, pf, P, horizon=30.0, seed=5)
        g64 = hjm.portfolio_cva_greeks(m, pf, P, horizon=30.0, precision=hjm.FP64, seed=5)
        assert g32.cva == pytest.approx(b.cva.sum(), rel=1e-12)
        assert g64.cva == pytest.approx(a.cva.sum(), rel=1e-12)
        for x, y in ((g32.d_spot, g64.d_spot), (g32.d_vol, g64.d_vol),
                     (g32.d_drift, g64.d_drift), (g32.d_vol_total, g64.d_vol_total)):
            assert np.all(np.isfinite(x))
            assert np.max(np.abs(x - y)) <= 1e-4 * np.max(np.abs(y))


def test_portfolio_greeks_match_oracle_adjoint(small, oracle):
    """128 paths, and 3 and 1 (a partly live workgroup of the seed and sweep kernels)."""
    m, pf, z, _ = small
    w = oracle.portfolio_weights(pf, 30.0)
    for P in (128, 3, 1):
        _flush("[gpu] portfolio adjoint vs oracle adjoint, %d paths x 30y, FP64 forward" % P)
        zp = np.ascontiguousarray(z[:P])
        cva, ds, dv, dm = oracle.portfolio_cva_greeks(m.spot_rates, m.drifts, m.volatilities,
                                                      30.0, 0.5, 0.01, zp, pf, w)
        g = hjm.portfolio_cva_greeks(m, pf, P, horizon=30.0, precision=hjm.FP64, normals=zp)
        assert g.cva == pytest.approx(cva, rel=1e-10)
        for got, ref in ((g.d_spot, ds), (g.d_vol, dv), (g.d_drift, dm)):
            assert np.max(np.abs(got - ref)) <= 1e-4 * np.max(np.abs(ref))


def _forward_only_ms(m, pf, paths, reps=3, **kw):
    """Device time of a forward-only portfolio plan (paths, valuation, statistics, no PFE):
    the median over a few executions."""
    plan = hjm.PortfolioPlan(m, pf, paths, horizon=30.0, pfe_quantile=0.0, **kw)
    plan.execute()
    t = float(np.median([plan.execute().timings_ms[4] for _ in range(reps)]))
    plan.close()
    return t


def test_portfolio_fp32_greeks_close_to_fp64_and_regenerated_shocks():
    """The default FP32 portfolio Greeks (matrix-core forward that keeps its shocks, one-pass
    reverse sweep reading them) against the FP64 forward's on the same in-kernel normals,
    65536 paths x 30y, 1024 swaps in 128 sets; and the sweep that regenerates the shocks
    instead gives the same Greeks."""
    _flush("[gpu] portfolio Greeks FP32 vs FP64, 65536 paths x 30y, 1024 swaps in 128 sets")
    m = hjm.benchmark_market()
    pf = hjm.synthetic_portfolio(1024, 128, seed=2024)
    a = hjm.portfolio_cva_greeks(m, pf, 65536, horizon=30.0, precision=hjm.FP64, seed=2024)
    b = hjm.portfolio_cva_greeks(m, pf, 65536, horizon=30.0, seed=2024)
    assert b.stored_normals_bytes == 65536 * 3000 * 3 * 4 and a.stored_normals_bytes == 0
    assert abs(b.cva / a.cva - 1) < 1e-4
    for x, y in ((b.d_spot, a.d_spot), (b.d_vol, a.d_vol), (b.d_drift, a.d_drift),
                 (b.d_vol_total, a.d_vol_total)):
        assert np.max(np.abs(x - y)) <= 1e-3 * np.max(np.abs(y))
    _flush("[gpu] portfolio Greeks FP32 with regenerated shocks")
    r = hjm.portfolio_cva_greeks(m, pf, 65536, horizon=30.0, seed=2024, regenerate_normals=True)
    assert r.stored_normals_bytes == 0 and r.cva == b.cva
    for x, y in ((r.d_spot, b.d_spot), (r.d_vol, b.d_vol), (r.d_drift, b.d_drift)):
        assert np.allclose(x, y, rtol=1e-12, atol=0)


def test_full_config_greeks_cost_ratio():
    _flush("[gpu] portfolio Greeks cost: 65536 paths x 30y, 1024 swaps in 128 sets, FP32")
    m = hjm.benchmark_market()
    pf = hjm.synthetic_portfolio(1024, 128, seed=2024)
    t = np.array([hjm.portfolio_cva_greeks(m, pf, 65536, horizon=30.0).timings_ms
                  for _ in range(3)])
    fwd = _forward_only_ms(m, pf, 65536)
    ratio = float(np.median(t[:, 2])) / fwd
    _flush("  forward-only %.2f ms, Greeks forward %.2f ms, adjoint %.2f ms, ratio %.2f"
           % (fwd, np.median(t[:, 0]), np.median(t[:, 1]), ratio))
    g = hjm.portfolio_cva_greeks(m, pf, 4096, horizon=30.0)
    assert g.cva > 0 and np.all(np.isfinite(g.d_vol_total))
    # the FP32 forward (path kernel and valuation) on the matrix cores, the seed and the
    # one-pass reverse sweep on the VALU; the target is 4
    assert ratio <= 4.0


def test_full_config_forward_and_greeks():
    """The full configuration on one GPU: 1048576 paths x 3000 steps (30y, dt = 0.01), 1024
    swaps in 128 netting sets (seed 2024). FP32 + compensated forward against the FP64
    forward on the same normals: total CVA within 1e-4, every set's CVA within 1e-3. The
    checkpointed adjoint in FP32 (its stored shocks reported) and in FP64: total CVA equal
    to the forward's, all 204 Greeks finite and FP32 within 1e-3 of FP64 (x scale)."""
    m = hjm.benchmark_market()
    pf = hjm.synthetic_portfolio(1024, 128, seed=2024)
    P = 1048576
    _flush("[gpu] full config: FP64 forward, %d paths x 30y, 1024 swaps in 128 sets" % P)
    a = hjm.simulate_portfolio(m, pf, P, horizon=30.0, precision=hjm.FP64, seed=2024)
    _flush("[gpu] full config: FP32 forward")
    b = hjm.simulate_portfolio(m, pf, P, horizon=30.0, precision=hjm.FP32_KAHAN, seed=2024)
    assert np.all(a.cva > 0) and np.all(a.cva_stderr < 0.1 * a.cva)
    assert abs(b.cva.sum() / a.cva.sum() - 1) < 1e-4
    assert np.max(np.abs(b.cva / a.cva - 1)) < 1e-3
    _flush("[gpu] full config: FP32 forward + adjoint")
    g32 = hjm.portfolio_cva_greeks(m, pf, P, horizon=30.0, seed=2024)
    _flush("  stored shocks %.1f GB, forward %.1f ms, adjoint %.1f ms"
           % (g32.stored_normals_bytes / 1e9, g32.timings_ms[0], g32.timings_ms[1]))
    assert g32.stored_normals_bytes in (0, P * 3000 * 3 * 4)
    _flush("[gpu] full config: FP64 forward + adjoint")
    g64 = hjm.portfolio_cva_greeks(m, pf, P, horizon=30.0, precision=hjm.FP64, seed=2024)
    assert g32.cva == pytest.approx(b.cva.sum(), rel=1e-12)
    assert g64.cva == pytest.approx(a.cva.sum(), rel=1e-12)
    for x, y in ((g32.d_spot, g64.d_spot), (g32.d_vol, g64.d_vol), (g32.d_drift, g64.d_drift),
                 (g32.d_vol_total, g64.d_vol_total)):
        assert np.all(np.isfinite(x)) and np.all(np.isfinite(y))
        assert np.max(np.abs(x - y)) <= 1e-3 * np.max(np.abs(y))
