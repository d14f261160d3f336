"""The command-line driver (successor of simulation_gpu_testing.cpp:175-259): flags, the
printed "%1.6f" expected-exposure profile, error exits.

CPU: argument errors and the no-device exit. GPU: the printed profile equals the library's
drop-in result (to the 6 printed decimals) on the benchmark market and on the reference
driver's own market and swap (tests/golden/legacy_market.json); -cva matches
simulate_reference.
"""
import json
import os
import subprocess

import numpy as np
import pytest

import cva_hjm_cuda_amd as hjm

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "cva_hjm_cuda_amd", "lib", "hjm_cli")


def _run(*args):
    if not os.path.exists(CLI):
        pytest.skip("driver not built")
    return subprocess.run([CLI] + [str(a) for a in args], capture_output=True, text=True,
                          timeout=300)


def _profile(out, title="Expected Exposure Profile Double"):
    lines = out.splitlines()
    i = lines.index(title)
    return np.array([float(x) for x in lines[i + 1].split()])


def _legacy_market_file(tmp_path):
    mk = json.load(open(os.path.join(ROOT, "tests", "golden", "legacy_market.json")))
    path = tmp_path / "legacy_market.txt"
    with open(path, "w") as f:
        for key in ("spot_rates", "volatilities", "drifts", "accrual"):
            f.write(key + ": " + " ".join(repr(float(v)) for v in mk[key]) + "\n")
    return path, mk


def test_cli_argument_errors(tmp_path):
    r = _run("-cpu")
    assert r.returncode == 2 and "GPU only" in r.stderr
    r = _run("-scenarios")
    assert r.returncode == 2 and "missing its value" in r.stderr
    r = _run("-dt", "abc")
    assert r.returncode == 2
    bad = tmp_path / "bad.txt"
    bad.write_text("spot_rates: 0.01 0.02\n")
    r = _run("-market", bad)
    assert r.returncode == 2 and "needs 51 spot_rates" in r.stderr
    r = _run("-market", tmp_path / "missing.txt")
    assert r.returncode == 2 and "cannot open" in r.stderr


def _devices():
    try:
        return hjm.device_count()
    except hjm.HJMError:
        return 0


def test_cli_without_device_fails_loudly():
    if _devices() > 0:
        pytest.skip("a GPU is visible")
    r = _run("-scenarios", 100)
    assert r.returncode == 1 and "calculateExposureMultiGPU failed" in r.stderr
    assert r.stdout.startswith("valid arguments")


@pytest.mark.gpu
def test_cli_profile_equals_dropin_benchmark_market():
    m = hjm.benchmark_market()
    swap = hjm.InterestRateSwap(1e6, 0.025, 10.0)
    r = _run("-scenarios", 4096, "-notional", 1e6, "-K", 0.025, "-expiry", 10)
    assert r.returncode == 0, r.stderr
    ee = np.zeros(51)
    hjm.calculate_exposure_multi_gpu(ee, swap, m.accrual, m.spot_rates, m.drifts,
                                     m.volatilities, 0.0, 1, 4096, 0.01)
    assert np.max(np.abs(_profile(r.stdout) - ee)) <= 5e-7 + 1e-12 * np.max(ee)


@pytest.mark.gpu
def test_cli_reference_driver_market_and_swap(tmp_path):
    path, mk = _legacy_market_file(tmp_path)
    r = _run("-scenarios", 5000, "-market", path, "-multi_gpu")
    assert r.returncode == 0, r.stderr
    m = hjm.Market(np.array(mk["spot_rates"]), np.array(mk["drifts"]),
                   np.array(mk["volatilities"]), np.array(mk["accrual"]))
    s = mk["swap"]
    ee = np.zeros(51)
    hjm.calculate_exposure_multi_gpu(ee, hjm.InterestRateSwap(s["notional"], s["K"], s["expiry"]),
                                     m.accrual, m.spot_rates, m.drifts, m.volatilities, 0.0, 4,
                                     5000, 0.01)
    got = _profile(r.stdout)
    assert np.max(np.abs(got - ee)) <= 5e-7
    assert got[50] == 0 and np.all(got >= 0)


@pytest.mark.gpu
def test_cli_cva_matches_simulate_reference():
    m = hjm.benchmark_market()
    r = _run("-cva", "-scenarios", 8192, "-notional", 1e6, "-K", 0.025, "-expiry", 10,
             "-seed", 77)
    assert r.returncode == 0, r.stderr
    ref = hjm.simulate_reference(m, hjm.InterestRateSwap(1e6, 0.025, 10.0), 8192, seed=77)
    cva_line = [l for l in r.stdout.splitlines() if l.startswith("CVA ")][0]
    assert float(cva_line.split()[1]) == pytest.approx(ref.cva, rel=1e-8)
    assert np.max(np.abs(_profile(r.stdout) - ref.expected_exposure)) <= 5e-7 + 1e-12 * 1e6
    assert np.max(np.abs(_profile(r.stdout, "PFE 97.5% Profile") - ref.pfe)) <= 5e-7 + 1e-6
