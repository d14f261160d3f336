"""Pins the normal-generator specification shared by the oracle and the product.

* Philox4x32-10 against the Random123 known-answer vectors and against rocRAND's
  open-source implementation (fixture tests/golden/philox_vectors.json).
* Box-Muller: the exact FP32 uniforms of a Philox pair, and the pair as the polar point
  (sqrt(-2 ln u), 2 pi t) to FP32 rounding, including the ends of both ranges.
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


def _uniforms(a, b):
    """The generator's exact FP32 uniforms: u = fma(float(a), 2^-32, 2^-33), t = float(b)
    2^-32 (float(x): round to nearest even; the fma's one rounding is that of an exact
    float64 sum narrowed to float32)."""
    fa = float(np.float32(np.uint32(a)))
    u = float(np.float32(fa * 2.0**-32 + 2.0**-33))
    t = float(np.float32(np.uint32(b))) * 2.0**-32
    return u, t


def test_box_muller_polar_form(oracle):
    """Each pair is the point of radius sqrt(-2 ln u) at angle 2 pi t, to FP32 rounding,
    for random words and the ends of both ranges."""
    rng = np.random.default_rng(3)
    words = [int(x) for x in rng.integers(0, 2**32, 6000)]
    edges = [0, 1, 127, 128, 129, 255, 256, 2**24 - 1, 2**24, 2**24 + 1, 2**31, 2**32 - 129,
             2**32 - 128, 2**32 - 1]
    pairs = list(zip(words[::2], words[1::2])) + [(a, b) for a in edges for b in edges]
    worst_r = worst_a = 0.0
    for a, b in pairs:
        z0, z1 = oracle.box_muller(a, b)
        u, t = _uniforms(a, b)
        assert 0.0 < u <= 1.0 and 0.0 <= t <= 1.0
        r = math.sqrt(-2.0 * math.log(u))
        worst_r = max(worst_r, abs(math.hypot(z0, z1) - r) / max(r, 1e-30))
        if r > 1e-3:
            d = math.atan2(z1, z0) - 2.0 * math.pi * t
            d = (d + math.pi) % (2.0 * math.pi) - math.pi
            worst_a = max(worst_a, abs(d))
    assert worst_r < 2.0**-23, worst_r      # each coordinate rounded once to FP32
    assert worst_a < 2.0**-21, worst_a
    # the ends of the radius: the smallest u is 2^-33 (r = sqrt(66 ln 2)); the largest
    # words round to u = 1 and give the origin
    z0, z1 = oracle.box_muller(0, 0)
    assert z1 == 0.0 and z0 == np.float32(math.sqrt(66.0 * math.log(2.0)))
    assert all(v == 0.0 for v in oracle.box_muller(2**32 - 1, 12345))
    # quarter turns: t = 1/4 puts the whole radius on the sine
    z0, z1 = oracle.box_muller(2**31, 2**30)
    assert abs(z0) < 1e-7 * z1 and z1 == np.float32(math.sqrt(-2.0 * math.log(0.5 + 2.0**-33)))


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
