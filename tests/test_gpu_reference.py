"""GPU parity of the reference-semantics path (HIP kernels through the C ABI) with the
oracle (restated simulation_cpu.h), on identical normals, plus size-independent
properties at the full benchmark size.

Tolerances: FP64 per-date EE within 1e-10 relative of the oracle (exactly 0 where the
oracle is 0); FP32 + compensated within 1e-4 relative on CVA; PFE is the exact order
statistic of the returned exposures. Normals: the GPU's Box-Muller (hardware log2, sqrt,
sin, cos on the exact Philox uniforms) is within 3.3e-7 |z| + 9e-7 of the oracle's FP64
transform; runs with in-kernel normals are compared with the oracle on the GPU's own
normals (hjm_generate_normals), which every path kernel regenerates bit for bit.
"""
import json
import math
import os

import numpy as np
import pytest

import cva_hjm_cuda_amd as hjm

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))

# Plain-method configuration: 1 payer IRS, N=1e6, K=2.5%, 5y, semiannual; 1000 paths of
# host-supplied normals (seed 42); flat hazard 2%, recovery 40%.
C1 = dict(paths=1000, expiry=5.0, notional=1e6, K=0.025, seed=42, hazard=0.02, recovery=0.4)


def _flush(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def c1(oracle):
    _flush("[gpu] config-1 oracle run")
    m = hjm.benchmark_market()
    S = int(C1["expiry"] / 0.01)
    z = oracle.normals(C1["seed"], C1["paths"], S)
    nx, ny = oracle.reference_paths(m.spot_rates, m.drifts, m.volatilities, C1["expiry"], 0.5,
                                    0.01, z)
    ex = oracle.reference_exposure(nx, ny, m.accrual, C1["notional"], C1["K"], 0.5)
    w = oracle.cva_weights(51, 0.5, C1["hazard"], C1["recovery"])
    ee, pfe, cva, se = oracle.stats(ex, w, 0.975)
    swap = hjm.InterestRateSwap(C1["notional"], C1["K"], C1["expiry"])
    return dict(m=m, z=z, nx=nx, ny=ny, ex=ex, ee=ee, pfe=pfe, cva=cva, se=se, swap=swap)


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    den = np.maximum(np.abs(b), 1e-300)
    return np.where(b == 0, np.abs(a), np.abs(a - b) / den)


def _normals_bound(g, o):
    """|GPU - oracle| over the bound of the hardware transform (3.3e-7 |z| + 9e-7): <= 1."""
    return np.max(np.abs(g - o) / (3.3e-7 * np.abs(o) + 9e-7))


def test_gpu_normals_within_bound_of_oracle(oracle):
    _flush("[gpu] normals")
    for seed, p0, stride in ((42, 0, 50), (7, 1000, 50), (0x0123456789ABCDEF, 123456, 50),
                             (42, 17, 20), (9, 5, 7)):
        g = hjm.generate_normals(seed, 48, 260, path_offset=p0, stride=stride)
        o = oracle.normals(seed, 48, 260, path0=p0, stride=stride)
        assert _normals_bound(g, o) <= 1.0, (seed, stride)
        # the same Philox words: every value is the oracle's to a few FP32 ulps
        assert np.median(np.abs(g - o)) < 1e-7, (seed, stride)


def _distinct_float_words():
    """Chunks of 32-bit words covering every distinct float32(word): all words below 2^24
    (exact), then each octave [2^k, 2^(k+1)), k = 24..31, at its float spacing 2^(k-23),
    and 2^32 - 1 (which rounds to 2^32)."""
    yield np.arange(0, 1 << 23, dtype=np.uint64)
    yield np.arange(1 << 23, 1 << 24, dtype=np.uint64)
    for k in range(24, 32):
        yield (1 << k) + (np.arange(1 << 23, dtype=np.uint64) << (k - 23))
    yield np.array([(1 << 32) - 1], dtype=np.uint64)


def test_box_muller_factors_within_their_bounds_exhaustive():
    """The hardware transcendentals of the generator's Box-Muller (hjm_rng.h) against the
    exact transform in FP64, over every distinct uniform: radius within 1.4e-7 relative,
    cosine and sine within 1.3e-7 absolute. These two bounds give the normals' bound
    3.3e-7 |z| + 9e-7 that the other normals tests apply (r <= 6.77)."""
    _flush("[gpu] Box-Muller factors, every distinct uniform (84 M words)")
    worst_r = worst_cs = 0.0
    for a in _distinct_float_words():
        r, c, s = hjm.box_muller_parts(a.astype(np.uint32))
        af = a.astype(np.float32).astype(np.float64)    # float(word), as the kernels round it
        u = (af * 2.0 ** -32 + 2.0 ** -33).astype(np.float32).astype(np.float64)  # one rounding
        re = np.sqrt(-2.0 * np.log(u))
        t = 2.0 * np.pi * (af * 2.0 ** -32)
        pos = re > 0
        assert np.all(r[~pos] == 0)
        if pos.any():
            worst_r = max(worst_r, float(np.max(np.abs(r[pos] - re[pos]) / re[pos])))
        worst_cs = max(worst_cs, float(np.max(np.abs(c - np.cos(t)))),
                       float(np.max(np.abs(s - np.sin(t)))))
    print("  radius max rel %.4e, cos/sin max abs %.4e" % (worst_r, worst_cs))
    assert worst_r <= 1.4e-7 and worst_cs <= 1.3e-7
    # the normals' bound from the factors': |r c - re ce| <= |r - re| |c| + re |c - ce|
    # plus the product's rounding (2^-24 relative), with re <= sqrt(66 ln 2)
    rmax = math.sqrt(66.0 * math.log(2.0))
    assert 1.4e-7 + 2.0 ** -24 <= 3.3e-7 and rmax * 1.3e-7 <= 9e-7


def test_gpu_normals_large_sample(oracle):
    """2048 paths x 1000 steps (6.1 M normals, 3 M Box-Muller pairs) against the oracle
    within the transform's bound; moments and a Kolmogorov-Smirnov test of the GPU's own
    normals; the three factors of a step uncorrelated."""
    from scipy import stats
    _flush("[gpu] normals, large sample")
    g = hjm.generate_normals(7, 2048, 1000, path_offset=99991)
    o = oracle.normals(7, 2048, 1000, path0=99991)
    assert _normals_bound(g, o) <= 1.0
    z = g.reshape(-1)
    n = z.size
    assert abs(z.mean()) < 4.0 / math.sqrt(n)
    assert abs(z.var() - 1.0) < 4.0 * math.sqrt(2.0 / n)
    assert abs(stats.skew(z)) < 4.0 * math.sqrt(6.0 / n)
    assert abs(stats.kurtosis(z)) < 4.0 * math.sqrt(24.0 / n)
    assert stats.kstest(z[:1000000], "norm").pvalue > 1e-3
    cc = np.corrcoef(g.reshape(-1, 3).T)
    assert np.max(np.abs(cc - np.eye(3))) < 4.0 / math.sqrt(n / 3)


def test_fp64_host_normals_matches_oracle(c1):
    _flush("[gpu] config-1 FP64 host normals")
    r = hjm.simulate_reference(c1["m"], c1["swap"], C1["paths"], normals=c1["z"],
                               hazard=C1["hazard"], recovery=C1["recovery"], keep_paths=True)
    D = 10
    # forwards are O(1e-2) and may cross zero: absolute check; discount factors relative
    assert np.max(np.abs(r.num_x[:, 1:D + 1] - c1["nx"][:, 1:D + 1])) < 1e-14
    assert np.max(_rel(r.num_y[:, 1:D + 1], c1["ny"][:, 1:D + 1])) < 1e-12
    scale = np.max(np.abs(c1["ex"]))
    assert np.max(np.abs(r.exposures - c1["ex"])) < 1e-10 * scale
    ee_o = c1["ee"]
    assert np.all(r.expected_exposure[ee_o == 0] == 0)
    assert np.max(_rel(r.expected_exposure[ee_o > 0], ee_o[ee_o > 0])) < 1e-10
    assert np.all(r.expected_exposure[D:] == 0)
    assert abs(r.cva / c1["cva"] - 1) < 1e-10
    assert abs(r.cva_stderr / c1["se"] - 1) < 1e-8
    pf = c1["pfe"]
    assert np.max(np.abs(r.pfe - pf)) <= 1e-10 * max(1.0, np.max(pf))


def test_fp64_in_kernel_rng_equals_host_normals(c1):
    _flush("[gpu] config-1 FP64 in-kernel philox vs host normals")
    z = hjm.generate_normals(C1["seed"], C1["paths"], int(C1["expiry"] / 0.01))
    a = hjm.simulate_reference(c1["m"], c1["swap"], C1["paths"], normals=z)
    b = hjm.simulate_reference(c1["m"], c1["swap"], C1["paths"], seed=C1["seed"])
    assert np.array_equal(a.expected_exposure, b.expected_exposure)
    assert a.cva == b.cva and np.array_equal(a.pfe, b.pfe)


def test_blockwise_baseline_matches_oracle(c1):
    _flush("[gpu] config-1 blockwise baseline")
    r = hjm.simulate_reference(c1["m"], c1["swap"], C1["paths"], normals=c1["z"], blockwise=True,
                               keep_paths=True)
    assert np.max(np.abs(r.num_x[:, 1:11] - c1["nx"][:, 1:11])) < 1e-15
    assert np.max(_rel(r.num_y[:, 1:11], c1["ny"][:, 1:11])) < 1e-13
    ee_o = c1["ee"]
    assert np.max(_rel(r.expected_exposure[ee_o > 0], ee_o[ee_o > 0])) < 1e-12
    # in-kernel normals (materialised by its own normals pass) against host-supplied copies
    # of the same generator's output
    z = hjm.generate_normals(C1["seed"], C1["paths"], int(C1["expiry"] / 0.01))
    r1 = hjm.simulate_reference(c1["m"], c1["swap"], C1["paths"], normals=z, blockwise=True)
    r2 = hjm.simulate_reference(c1["m"], c1["swap"], C1["paths"], seed=C1["seed"], blockwise=True)
    assert np.allclose(r2.expected_exposure, r1.expected_exposure, rtol=1e-14, atol=0)


def test_legacy_fixture_fp64_matches_oracle(oracle):
    """The reference driver's own market data and swap (25y, N=10, K=6.7%), 2000 paths."""
    _flush("[gpu] legacy fixture 25y FP64")
    mk = json.load(open(os.path.join(HERE, "golden", "legacy_market.json")))
    m = hjm.Market(mk["spot_rates"], mk["drifts"], mk["volatilities"], mk["accrual"])
    P, S = 2000, 2500
    z = oracle.normals(1234, P, S)
    _, nx, ny = oracle.time_reference_paths(m.spot_rates, m.drifts, m.volatilities, 25.0, 0.5,
                                            0.01, z, nthreads=8)
    ex = oracle.reference_exposure(nx, ny, m.accrual, 10.0, 0.067, 0.5)
    ee = oracle.expected_exposure(ex)
    r = hjm.simulate_reference(m, hjm.InterestRateSwap(10.0, 0.067, 25.0), P, normals=z)
    assert ee[50] == 0 and r.expected_exposure[50] == 0
    assert np.max(_rel(r.expected_exposure[:50], ee[:50])) < 1e-10


def test_fp32_compensated_error_vs_fp64(c1):
    _flush("[gpu] config-1 FP32 compensated vs FP64")
    a = hjm.simulate_reference(c1["m"], c1["swap"], C1["paths"], normals=c1["z"])
    b = hjm.simulate_reference(c1["m"], c1["swap"], C1["paths"], normals=c1["z"],
                               precision=hjm.FP32_KAHAN)
    big = a.expected_exposure > 1e-3 * a.expected_exposure.max()
    assert np.max(_rel(b.expected_exposure[big], a.expected_exposure[big])) < 1e-4
    assert abs(b.cva / a.cva - 1) < 1e-4


def test_dropin_sharding_invariance(c1):
    _flush("[gpu] drop-in calculate_exposure_multi_gpu shards")
    m = c1["m"]
    swap = hjm.InterestRateSwap(1e6, 0.025, 5.0)
    outs = []
    for g in (1, 3, 4):
        ee = np.zeros(51)
        hjm.calculate_exposure_multi_gpu(ee, swap, m.accrual, m.spot_rates, m.drifts,
                                         m.volatilities, 0.0, g, 10001, 0.01)
        outs.append(ee)
    for ee in outs[1:]:
        assert np.max(_rel(ee[outs[0] > 0], outs[0][outs[0] > 0])) < 1e-13
    ref = hjm.simulate_reference(m, swap, 10001, seed=1234)
    assert np.max(_rel(outs[0][outs[0] > 0], ref.expected_exposure[outs[0] > 0])) < 1e-13
    ee32 = np.zeros(51, np.float32)
    hjm.calculate_exposure_multi_gpu(ee32, swap, m.accrual, m.spot_rates, m.drifts,
                                     m.volatilities, 0.0, 2, 10001, 0.01)
    big = outs[0] > 1e-3 * outs[0].max()
    assert np.max(_rel(ee32[big].astype(np.float64), outs[0][big])) < 1e-4


def test_cva_within_3_se_of_independent_oracle_run(oracle):
    """In-kernel RNG: the GPU CVA and an oracle CVA on independent normals agree within
    3 combined Monte Carlo standard errors (10y payer, 32768 GPU / 2048 oracle paths)."""
    _flush("[gpu] CVA statistical check vs independent oracle run")
    m = hjm.benchmark_market()
    swap = hjm.InterestRateSwap(1e6, 0.025, 10.0)
    g = hjm.simulate_reference(m, swap, 32768, seed=7)
    z = oracle.normals(8, 2048, 1000)
    _, nx, ny = oracle.time_reference_paths(m.spot_rates, m.drifts, m.volatilities, 10.0, 0.5,
                                            0.01, z, nthreads=8)
    ex = oracle.reference_exposure(nx, ny, m.accrual, 1e6, 0.025, 0.5)
    _, _, cva, se = oracle.stats(ex, oracle.cva_weights(51, 0.5, 0.02, 0.4), 0.975)
    assert abs(g.cva - cva) < 3.0 * math.sqrt(g.cva_stderr ** 2 + se ** 2)


@pytest.fixture(scope="module")
def full_size():
    _flush("[gpu] precision config 131072 x 1000, FP64 and FP32")
    m = hjm.benchmark_market()
    swap = hjm.InterestRateSwap(1e6, 0.025, 10.0)
    a = hjm.simulate_reference(m, swap, 131072, seed=7)
    b = hjm.simulate_reference(m, swap, 131072, seed=7, precision=hjm.FP32_KAHAN)
    return m, swap, a, b


def test_full_size_properties(full_size):
    m, swap, a, b = full_size
    for r in (a, b):
        assert np.all(r.expected_exposure >= 0) and np.all(r.expected_exposure[20:] == 0)
        assert np.all(r.pfe[:20] >= r.expected_exposure[:20] - 1e-9)
        assert r.cva > 0 and 0 < r.cva_stderr < 0.05 * r.cva
    assert abs(b.cva / a.cva - 1) < 1e-4
    big = a.expected_exposure > 1e-3 * a.expected_exposure.max()
    assert np.max(_rel(b.expected_exposure[big], a.expected_exposure[big])) < 1e-4


def test_full_size_deterministic_and_shard_additive(full_size):
    m, swap, a, _ = full_size
    a2 = hjm.simulate_reference(m, swap, 131072, seed=7)
    assert np.array_equal(a.sums, a2.sums) and np.array_equal(a.pfe, a2.pfe)
    h1 = hjm.simulate_reference(m, swap, 65536, seed=7, path_offset=0)
    h2 = hjm.simulate_reference(m, swap, 65536, seed=7, path_offset=65536)
    s = h1.sums + h2.sums
    assert np.max(_rel(s[a.sums > 0], a.sums[a.sums > 0])) < 1e-12


def test_pfe_is_exact_order_statistic(c1):
    _flush("[gpu] PFE exact rank")
    for P in (1000, 777, 1):
        r = hjm.simulate_reference(c1["m"], c1["swap"], P, seed=3, keep_paths=True)
        k = max(1, min(P, math.ceil(0.975 * P)))
        for t in range(10):
            col = r.exposures[:, t]
            assert r.pfe[t] == np.partition(col, k - 1)[k - 1]


def test_tiny_path_counts(c1, oracle):
    """1, 2 and 33 paths (a partly live wave, N-tile and workgroup in every pass): FP64 on
    host normals equals the oracle per path and in EE, PFE, CVA and its standard error (0
    for one path, as the oracle defines it); the FP32 matrix-core kernel's numeraires of
    those paths equal the first paths of a 64-path run bit for bit."""
    m, swap = c1["m"], c1["swap"]
    w = oracle.cva_weights(51, 0.5, C1["hazard"], C1["recovery"])
    scale = np.max(np.abs(c1["ex"]))
    for P in (1, 2, 33):
        _flush("[gpu] %d paths, FP64 host normals vs oracle" % P)
        r = hjm.simulate_reference(m, swap, P, normals=c1["z"][:P], hazard=C1["hazard"],
                                   recovery=C1["recovery"], keep_paths=True)
        ex = c1["ex"][:P]
        assert np.max(np.abs(r.exposures - ex)) < 1e-10 * scale
        ee, pfe, cva, se = oracle.stats(ex, w, 0.975)
        assert np.all(r.expected_exposure[ee == 0] == 0)
        assert np.max(_rel(r.expected_exposure[ee > 0], ee[ee > 0])) < 1e-10
        assert np.max(np.abs(r.pfe - pfe)) <= 1e-10 * scale
        assert abs(r.cva - cva) <= 1e-10 * abs(cva)
        if P == 1:
            assert r.cva_stderr == 0.0 and se == 0.0
        else:
            assert abs(r.cva_stderr - se) <= 1e-8 * se + 1e-10 * abs(cva)
    _flush("[gpu] 1, 2, 33 paths, FP32 matrix-core kernel vs the first paths of 64")
    big = hjm.simulate_reference(m, swap, 64, seed=9, precision=hjm.FP32_KAHAN, keep_paths=True,
                                 path_kernel=hjm.KERNEL_MATRIX)
    for P in (1, 2, 33):
        s = hjm.simulate_reference(m, swap, P, seed=9, precision=hjm.FP32_KAHAN, keep_paths=True,
                                   path_kernel=hjm.KERNEL_MATRIX)
        assert np.array_equal(s.num_x, big.num_x[:P]) and np.array_equal(s.num_y, big.num_y[:P])
        assert np.array_equal(s.exposures, big.exposures[:P])


def test_gpu_argument_errors(c1):
    with pytest.raises(hjm.HJMError) as ei:
        hjm.simulate_reference(c1["m"], c1["swap"], 64, device=99)
    assert ei.value.code == 2
    with pytest.raises(hjm.HJMError) as ei:
        hjm.simulate_reference(c1["m"], hjm.InterestRateSwap(1e6, 0.025, 30.0), 64)
    assert ei.value.code == 3


def test_select_kth_exact_on_hard_rows():
    """The PFE selection on crafted rows: ties, all zeros, a rank inside a bin too large
    for the LDS candidate buffer, extreme ranks, values spanning many binades."""
    _flush("[gpu] radix selection on crafted rows")
    rng = np.random.default_rng(12)
    n = 200003
    rows = [rng.lognormal(0.0, 3.0, n),                 # many binades
            np.zeros(n),                                # all equal (zero)
            np.where(rng.random(n) < 0.9, 0.0, rng.random(n)),
            np.full(n, 1234.5),                         # one value, > 4096 copies per bin
            np.round(rng.random(n) * 50.0) / 8.0,       # heavy ties
            1.0 + rng.random(n) * 1e-12,                # one binade, narrow mantissas
            np.sort(rng.random(n)),                     # leading sample unrepresentative
            np.sort(rng.random(n))[::-1].copy(),
            # representative leading sample, then a bulk concentrated at the quantile
            # (more candidates than the bracket's buffer holds)
            np.concatenate([rng.random(8192), 0.975 + 1e-9 * rng.random(n - 8192)])]
    X = np.stack(rows)
    for k in (1, 2, 4097, math.ceil(0.975 * n), n - 1, n):
        got = hjm.select_kth(X, k)
        want = np.partition(X, k - 1, axis=1)[:, k - 1]
        assert np.array_equal(got, want), k
    # even row length (aligned vector loads in the counting pass) and random ranks
    Y = rng.lognormal(0.0, 1.0, (5, 1 << 17)) * (rng.random((5, 1 << 17)) < 0.6)
    for k in rng.integers(1, 1 << 17, 6):
        got = hjm.select_kth(Y, int(k))
        assert np.array_equal(got, np.partition(Y, k - 1, axis=1)[:, k - 1]), k
    with pytest.raises(hjm.HJMError):
        hjm.select_kth(-X[:1, :10], 3)
    with pytest.raises(hjm.HJMError):
        hjm.select_kth(X[:1, :10], 11)


def test_in_kernel_normals_at_another_stride(oracle):
    """dt = 0.02, 25 steps per date: each block draws one whole 16-step group of class-stream
    shocks through LDS and 9 steps directly. FP64 (one path per lane) and the FP32 VALU
    kernel against the oracle on the generator's class-stream normals for that stride, and
    the adjoint's regenerated shocks against the same normals supplied from the host."""
    _flush("[gpu] in-kernel normals at stride 25 (dt = 0.02)")
    m = hjm.benchmark_market()
    swap = hjm.InterestRateSwap(1e6, 0.025, 5.0)
    P, S = 512, 250
    z = hjm.generate_normals(3, P, S, stride=25)
    assert _normals_bound(z, oracle.normals(3, P, S, stride=25)) <= 1.0
    nx, ny = oracle.reference_paths(m.spot_rates, m.drifts, m.volatilities, 5.0, 0.5, 0.02, z)
    ee = oracle.expected_exposure(oracle.reference_exposure(nx, ny, m.accrual, 1e6, 0.025, 0.5))
    pos = ee > 0
    r = hjm.simulate_reference(m, swap, P, dt=0.02, seed=3, precision=hjm.FP64)
    assert np.max(np.abs(r.expected_exposure[pos] / ee[pos] - 1)) < 1e-10
    r32 = hjm.simulate_reference(m, swap, P, dt=0.02, seed=3, precision=hjm.FP32_KAHAN)
    assert np.max(np.abs(r32.expected_exposure[pos] / ee[pos] - 1)) < 1e-4
    a = hjm.cva_greeks(m, swap, P, dt=0.02, seed=3, precision=hjm.FP64)
    b = hjm.cva_greeks(m, swap, P, dt=0.02, normals=z, precision=hjm.FP64)
    assert a.cva == pytest.approx(b.cva, rel=1e-12)
    for x, y in ((a.d_spot, b.d_spot), (a.d_vol, b.d_vol)):
        assert np.max(np.abs(x - y)) <= 1e-10 * np.max(np.abs(y))
