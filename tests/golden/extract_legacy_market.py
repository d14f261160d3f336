"""Extract the market-data arrays hard-coded in the reference's driver as a JSON fixture.

The reference's only executable check (``simulation_gpu_testing.cpp``) feeds its CPU and
GPU paths with literal arrays: the initial forward curve ``spot_rates[51]`` (:36-41), the
factor-major volatility loadings ``volatilities[153]`` (:47-63), the drift table
``drifts[51]`` (:66-71) and the 30/360 ``accrual`` fractions (:117-123), plus the swap
used by ``main`` (:147-148, :215: notional 10, K 6.7%, expiry 25y, dtau 0.5) and the
default ``dt`` (:227). This script parses those numbers out of the file's text and writes
``legacy_market.json`` beside itself. Only numbers travel; no source text is kept.

Usage:  python tests/golden/extract_legacy_market.py [path-to-simulation_gpu_testing.cpp]
"""
import json
import re
import sys
from pathlib import Path

DEFAULT_SRC = "/root/reference/simulation_gpu_testing.cpp"


def _numbers(block: str):
    return [float(x) for x in re.findall(r"-?\d+\.\d*(?:[eE][-+]?\d+)?|-?\d+(?:[eE][-+]?\d+)?", block)]


def _array(text: str, decl: str):
    m = re.search(re.escape(decl) + r"[^=]*=\s*\{(.*?)\};", text, re.S)
    if not m:
        raise KeyError(decl)
    return _numbers(m.group(1))


def main(argv):
    src = Path(argv[1] if len(argv) > 1 else DEFAULT_SRC).read_text()
    out = {
        "source": "simulation_gpu_testing.cpp",
        "spot_rates": _array(src, "real spot_rates[51]"),
        "volatilities": _array(src, "real volatilities[153]"),
        "drifts": _array(src, "real drifts[51]"),
        "accrual": _array(src, "std::vector<real> accrual"),
        "swap": {"notional": 10.0, "K": 0.067, "expiry": 25.0, "dtau": 0.5},
        "dt": 0.01,
        "rng_seed": 1234,
    }
    assert len(out["spot_rates"]) == 51 and len(out["drifts"]) == 51
    assert len(out["volatilities"]) == 153 and len(out["accrual"]) == 51
    dst = Path(__file__).with_name("legacy_market.json")
    dst.write_text(json.dumps(out, indent=1))
    print("wrote", dst)


if __name__ == "__main__":
    main(sys.argv)
