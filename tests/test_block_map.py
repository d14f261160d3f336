"""The matrix-core kernels' 50-step block maps (host side, no GPU).

Reference semantics: W maps [f_in; phi; 1] to [f after 50 Euler steps; sum of the 50
curves after each step], in the K order the kernel feeds the matrix cores. Pinned two ways:
against 50 explicit FP64 Euler steps of the Musiela SDE written out here (random curve and
shocks), and chained over the blocks of a 5-year run against the oracle's reference paths
(num_x, num_y) on the same normals. Portfolio semantics: [f after 50 steps; left-point sum
of the short rate f(., 0) over the 50 steps], against the same explicit steps.
"""
import numpy as np
import pytest

import cva_hjm_cuda_amd as hjm

TENORS, STRIDE, QUADS, K = 51, 50, 13, 224


def _k_vector(f_in, phi):
    """Input column in the kernel's K order: index 32 c + 8 g + i (chunk c, lane group g,
    slot i); chunks 0, 1 carry curve rows 16 (2c + i/4) + 4 g + i%4, chunks 2..6 carry value
    v = 8 (c - 2) + i of lane group g: factor v % 3 of step 4 (v // 3) + g, the constant 1 at
    step 50 factor 0."""
    x = np.zeros(K)
    for k in range(K):
        c, g, i = k // 32, (k % 32) // 8, k % 8
        if c < 2:
            R = 16 * (2 * c + i // 4) + 4 * g + i % 4
            if R < TENORS:
                x[k] = f_in[R]
        else:
            v = 8 * (c - 2) + i
            if v < 3 * QUADS:
                q, fac = divmod(v, 3)
                t = 4 * q + g
                if t < STRIDE:
                    x[k] = phi[t, fac]
                elif t == STRIDE and fac == 0:
                    x[k] = 1.0
    return x


def _euler(f, drift, vols, phi, dt, dtau, left=None):
    area = np.zeros(TENORS)
    for t in range(STRIDE):
        if left is not None:
            left[0] += f[0]
        d = np.empty(TENORS)
        d[:-1] = f[1:] - f[:-1]
        d[-1] = f[-1] - f[-2]
        f = f + drift * dt + np.sqrt(dt) * (vols.T @ phi[t]) + d * dt / dtau
        area += f
    return f, area


@pytest.fixture(scope="module")
def market():
    return hjm.benchmark_market()


def test_block_map_equals_fifty_euler_steps(market):
    W = hjm.block_map(market).astype(np.float64)
    assert W.shape == (102, 224)
    rng = np.random.default_rng(5)
    vols = market.volatilities.reshape(3, TENORS)
    for _ in range(3):
        f_in = 0.03 + 0.01 * rng.standard_normal(TENORS)
        phi = rng.standard_normal((STRIDE, 3))
        f_out, area = _euler(f_in, market.drifts, vols, phi, 0.01, 0.5)
        y = W @ _k_vector(f_in, phi)
        np.testing.assert_allclose(y[:TENORS], f_out, rtol=0, atol=2e-8)
        np.testing.assert_allclose(y[TENORS:], area, rtol=2e-7, atol=0)
    # every curve row, every (step, factor) and the constant appear exactly once in K;
    # the remaining columns are zero
    x = _k_vector(np.arange(TENORS) + 2.0, 1000.0 + np.arange(3 * STRIDE).reshape(STRIDE, 3))
    used = x[x != 0]
    assert len(used) == len(set(used)) == TENORS + 3 * STRIDE + 1
    assert not np.any(W[:, x == 0])


def test_block_map_chain_matches_oracle_paths(market, oracle):
    expiry, dtau, dt, P = 5.0, 0.5, 0.01, 6
    z = oracle.normals(17, P, int(round(expiry / dt)))
    nx, ny = oracle.reference_paths(market.spot_rates, market.drifts, market.volatilities,
                                    expiry, dtau, dt, z)
    W = hjm.block_map(market).astype(np.float64)
    D = int(round(expiry / dtau))
    for p in range(P):
        f = market.spot_rates.copy()
        A = STRIDE * market.spot_rates.copy()  # block 0 holds the spot curve
        for b in range(1, D):
            y = W @ _k_vector(f, z[p, b * STRIDE:(b + 1) * STRIDE])
            f, A = y[:TENORS], A + y[TENORS:]
            assert abs(f[b + 1] - nx[p, b + 1]) < 5e-8, (p, b)
            assert abs(np.exp(-A[b + 1] * dt) / ny[p, b + 1] - 1) < 1e-7, (p, b)


def test_portfolio_block_map_equals_fifty_euler_steps(market):
    W = hjm.block_map(market, portfolio=True).astype(np.float64)
    assert W.shape == (52, 224)
    rng = np.random.default_rng(9)
    vols = market.volatilities.reshape(3, TENORS)
    for _ in range(3):
        f_in = 0.03 + 0.01 * rng.standard_normal(TENORS)
        phi = rng.standard_normal((STRIDE, 3))
        left = [0.0]
        f_out, _ = _euler(f_in, market.drifts, vols, phi, 0.01, 0.5, left)
        y = W @ _k_vector(f_in, phi)
        np.testing.assert_allclose(y[:TENORS], f_out, rtol=0, atol=2e-8)
        assert abs(y[TENORS] / left[0] - 1) < 2e-7
    # the curve rows are the reference map's
    Wr = hjm.block_map(market).astype(np.float64)
    assert np.array_equal(W[:TENORS], Wr[:TENORS])
    # a 30y chain (60 blocks) against explicit steps: curve and log bank account
    f, fe, lnb, lnb_e = market.spot_rates.copy(), market.spot_rates.copy(), 0.0, [0.0]
    for b in range(60):
        phi = rng.standard_normal((STRIDE, 3))
        y = W @ _k_vector(f, phi)
        f, lnb = y[:TENORS], lnb + 0.01 * y[TENORS]
        fe, _ = _euler(fe, market.drifts, vols, phi, 0.01, 0.5, lnb_e)
    assert np.max(np.abs(f - fe)) < 1e-7
    assert abs(lnb / (0.01 * lnb_e[0]) - 1) < 1e-6


def test_block_map_rejects_other_grids(market):
    with pytest.raises(hjm.HJMError):
        hjm.block_map(market, dt=0.02)
    with pytest.raises(hjm.HJMError):
        hjm.block_map(market, dt=0.02, portfolio=True)


def _curve_column(k):
    """Curve row fed by K index k (chunks 0, 1), or None."""
    c, g, i = k // 32, (k % 32) // 8, k % 8
    if c >= 2:
        return None
    R = 16 * (2 * c + i // 4) + 4 * g + i % 4
    return R if R < TENORS else None


@pytest.mark.parametrize("portfolio", [False, True])
def test_block_maps_vanish_below_each_rows_first_tenor(market, portfolio):
    """The kernels skip the products whose W entries are structurally zero. T couples tenor
    j only to j + 1, and tenor 50 to 49 (the downwind extrapolation), so T^n -- and the
    integral rows' sums of powers -- have no entry in column R < min(j, 49) of row j; the
    integral row 51 + j follows tenor j, the portfolio's short-rate row 51 spans every
    tenor. Exactly zero below that bound, nonzero at it (the structure is not a
    coincidence of this market)."""
    W = hjm.block_map(market, portfolio=portfolio).astype(np.float64)
    for row in range(W.shape[0]):
        if row < TENORS:
            first = min(row, 49)
        else:
            first = 0 if portfolio else min(row - TENORS, 49)
        for k in range(K):
            R = _curve_column(k)
            if R is None:
                continue
            if R < first:
                assert W[row, k] == 0.0, (row, R)
            elif R == first:
                assert W[row, k] != 0.0, (row, R)


def _mfma64_block(frag, f_in, phi):
    """One block as the FP64 matrix-core kernels compute it from their A operands: per
    M-tile the sum over the 52 instructions of A_j (16 x 4, lane l holding A[l % 16][l // 16])
    times B_j (4 x 16, lane l holding B[l // 16][l % 16]), lane group g supplying what the
    kernel's registers hold: curve row 16 (j // 4) + 4 (j % 4) + g for j < 13 (0 past row
    50), factor v % 3 of step 4 (v // 3) + g for v = j - 13 (1 at step 50 factor 0, 0
    past it). One path: every column of B is the same."""
    mt = frag.shape[0]
    out = np.zeros(16 * mt)
    for m in range(mt):
        for j in range(52):
            A = np.zeros((16, 4))
            for lane in range(64):
                A[lane % 16, lane // 16] = frag[m, j, lane]
            B = np.zeros(4)
            for g in range(4):
                if j < 13:
                    R = 16 * (j // 4) + 4 * (j % 4) + g
                    B[g] = f_in[R] if R < TENORS else 0.0
                else:
                    q, fac = divmod(j - 13, 3)
                    t = 4 * q + g
                    B[g] = phi[t, fac] if t < STRIDE else float(t == STRIDE and fac == 0)
            out[16 * m:16 * m + 16] += A @ B
    return out


@pytest.mark.parametrize("portfolio", [False, True])
def test_fp64_block_map_operands_equal_fifty_euler_steps(market, portfolio):
    """The FP64 kernels' A operands (hjm_block_map64), fed the B operands their registers
    hold, give the block's 50 FP64 Euler steps: the curve, and the integral increments
    (reference) or the short rate's left-point sum (portfolio), to FP64 rounding."""
    frag = hjm.block_map64(market, portfolio=portfolio)
    assert frag.shape == ((4 if portfolio else 7), 52, 64)
    rng = np.random.default_rng(13)
    vols = market.volatilities.reshape(3, TENORS)
    for _ in range(2):
        f_in = 0.03 + 0.01 * rng.standard_normal(TENORS)
        phi = rng.standard_normal((STRIDE, 3))
        left = [0.0]
        f_out, area = _euler(f_in, market.drifts, vols, phi, 0.01, 0.5, left)
        y = _mfma64_block(frag, f_in, phi)
        np.testing.assert_allclose(y[:TENORS], f_out, rtol=1e-12, atol=1e-15)
        if portfolio:
            assert abs(y[TENORS] / left[0] - 1) < 1e-12
            assert not np.any(y[TENORS + 1:])
        else:
            np.testing.assert_allclose(y[TENORS:2 * TENORS], area, rtol=1e-12, atol=0)
            assert not np.any(y[2 * TENORS:])
    # the structural zeros the kernels skip are zeros of these operands
    for m in range(frag.shape[0]):
        for j in range(13):
            if 16 * (j // 4) + 4 * (j % 4) + 3 < _tile_first_input(m, portfolio):
                assert not np.any(frag[m, j]), (m, j)
    with pytest.raises(hjm.HJMError):
        hjm.block_map64(market, portfolio=portfolio, mtiles=8)


def _tile_first_input(m, portfolio):
    rows = TENORS + 1 if portfolio else 2 * TENORS
    first = []
    for row in range(16 * m, min(16 * m + 16, rows)):
        if row < TENORS:
            first.append(min(row, 49))
        else:
            first.append(0 if portfolio else min(row - TENORS, 49))
    return min(first) if first else TENORS
