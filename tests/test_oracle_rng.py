"""Pins the normal-generator specification shared by the oracle and the product.

* Philox4x32-10 against the Random123 known-answer vectors and against rocRAND's
  open-source implementation (fixture tests/golden/philox_vectors.json).
* The FP32 polynomial ln / sin / cos against float64 numpy, in float32 ulp.
* The normals: moments and a Kolmogorov-Smirnov test on a large sample.
* The class-stream layout, restated from the Philox and Box-Muller primitives.
"""
import json
import math
import os

import numpy as np
import pytest
from scipy import stats

HERE = os.path.dirname(os.path.abspath(__file__))

# Random123 kat_vectors, philox4x32 with 10 rounds: (ctr0..3, key0..1) -> out0..3.
RANDOM123_KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,out", RANDOM123_KAT)
def test_philox_random123_kat(oracle, ctr, key, out):
    assert tuple(int(v) for v in oracle.philox(ctr, key)) == out


def test_philox_matches_rocrand_vectors(oracle):
    with open(os.path.join(HERE, "golden", "philox_vectors.json")) as f:
        vecs = json.load(f)["vectors"]
    assert len(vecs) >= 64
    for v in vecs:
        assert list(int(x) for x in oracle.philox(v["ctr"], v["key"])) == v["out"]


def _ulp(x):
    return np.spacing(np.float32(abs(x))).astype(np.float64)


def test_log_f32_accuracy(oracle):
    rng = np.random.default_rng(1)
    ks = np.concatenate([np.arange(1, 4000), rng.integers(1, 2**24 + 1, 20000), [2**24]])
    worst = 0.0
    for k in ks:
        u = np.float32(k * 2.0**-24)
        got = oracle.log_f32(u)
        ref = math.log(float(u))
        if ref == 0.0:
            assert got == 0.0
            continue
        worst = max(worst, abs(got - ref) / _ulp(ref))
    assert worst <= 2.0, worst


def test_sincos_turn24_accuracy(oracle):
    rng = np.random.default_rng(2)
    ks = np.concatenate([np.arange(0, 2**24, 4099), rng.integers(0, 2**24, 20000),
                         [0, 2**22, 2**23, 3 * 2**22, 2**21, 2**24 - 1]])
    worst = 0.0
    for k in ks:
        s, c = oracle.sincos_turn24(int(k))
        a = 2 * math.pi * int(k) * 2.0**-24
        for got, ref in ((s, math.sin(a)), (c, math.cos(a))):
            # absolute error in units of the ulp of 1/8 near zero crossings
            worst = max(worst, abs(got - ref) / max(_ulp(ref), 2.0**-27))
    assert worst <= 3.0, worst
    # exact quarter turns
    assert oracle.sincos_turn24(0) == (0.0, 1.0)
    assert oracle.sincos_turn24(2**22) == (1.0, -0.0) or oracle.sincos_turn24(2**22)[0] == 1.0


def test_normals_distribution(oracle):
    z = oracle.normals(seed=7, npaths=400, nsteps=1000).reshape(-1)
    n = z.size
    assert abs(z.mean()) < 4.0 / math.sqrt(n)
    assert abs(z.var() - 1.0) < 4.0 * math.sqrt(2.0 / n)
    assert abs(stats.skew(z)) < 4.0 * math.sqrt(6.0 / n)
    assert abs(stats.kurtosis(z)) < 4.0 * math.sqrt(24.0 / n)
    assert stats.kstest(z, "norm").pvalue > 1e-3
    # the three factors of a step are uncorrelated
    f = oracle.normals(seed=11, npaths=200, nsteps=1000).reshape(-1, 3)
    cc = np.corrcoef(f.T)
    assert np.max(np.abs(cc - np.eye(3))) < 4.0 / math.sqrt(f.shape[0])


def test_normals_counter_addressing(oracle):
    """Step n of path p depends only on (seed, p, n): any sub-range reproduces the whole."""
    whole = oracle.normals(seed=42, npaths=9, nsteps=60)
    part = oracle.normals(seed=42, npaths=4, nsteps=60, path0=5)
    assert np.array_equal(whole[5:], part)
    z = oracle.step_normals(42, 7, 33)
    assert np.array_equal(whole[7, 33], z.astype(np.float64))
    other = oracle.normals(seed=43, npaths=9, nsteps=60)
    assert not np.any(whole == other)


def test_class_stream_layout(oracle):
    """Step n of a run with date blocks of S steps: block b = n // S, offset s = n % S,
    class g = s % 4, q = s // 4; factor k is shock j = 3 q + k of the stream (b, g), i.e.
    element j % 4 of the Philox block with counter (b, path, g, j // 4), whose outputs go
    through Box-Muller as the pairs (x0, x1) and (x2, x3). Restated here from the
    primitives, for several strides (the product's kernels follow the same layout)."""
    seed = (7 << 32) | 12345
    key = (seed & 0xFFFFFFFF, seed >> 32)
    for stride in (50, 20, 7, 4, 1):
        for path in (0, 3, 99991):
            for n in range(3 * stride + 5):
                b, s = divmod(n, stride)
                g, q = s % 4, s // 4
                want = []
                for k in range(3):
                    j = 3 * q + k
                    x = [int(v) for v in oracle.philox((b, path, g, j // 4), key)]
                    z = oracle.box_muller(x[0], x[1]) + oracle.box_muller(x[2], x[3])
                    want.append(z[j % 4])
                got = oracle.step_normals(seed, path, n, stride)
                assert np.array_equal(got, np.array(want, np.float32)), (stride, path, n)
    # the stream-walking fill (each call once, shocks scattered) equals the step form
    for stride in (50, 20, 7, 4, 1):
        z = oracle.normals(seed, 3, 2 * stride + 3, path0=11, stride=stride)
        for n in range(z.shape[1]):
            assert np.array_equal(z[1, n], oracle.step_normals(seed, 12, n, stride)), (stride, n)
    # one call of a 50-step block feeds four shocks: 38 calls cover its 150 shocks
    calls = {(g, (3 * (s // 4) + k) // 4) for s in range(50) for k in range(3) for g in [s % 4]}
    assert len(calls) == 38
