"""Benchmark of the HJM path-generation + swap-exposure hot path on one MI355X.

Workload (default): the precision configuration -- 10y payer IRS (N=1e6, K=2.5%),
Nelson-Siegel curve, 3 parametric factors, 51 tenors, dt=0.01 -> 1000 steps,
131072 paths, in-kernel Philox normals (seed 7), FP32 evolution with Kahan-compensated
integrals, EE / PFE97.5 / CVA (h=2%, R=40%) in FP64. One step = one full pass (path
kernel, exposure, reductions, PFE selection) with inputs resident on the device.

Prints one JSON line. Also measured and reported: the FP64 kernels (vector, and matrix
cores: `fp64_matrix`) on the same normals (path-steps/s and the FP32-vs-FP64 EE/PFE/CVA
errors), the reference's own GPU
design in FP64 ("plain FP64 GPU form", one timed pass), and the CPU oracle on a bounded
sample (cpu_baseline).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--config precision|batched|full]

--config batched / full run the portfolio workloads instead (256 swaps in 32 netting sets
x 262144 paths, resp. 1024 swaps in 128 sets x 1048576 paths, 30y, FP32 + compensated,
launch shapes of the path and valuation kernels chosen by the autotuner before the timed
region, one hipGraph replay per step).
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FLOPS_PER_PATH_STEP = 561          # 51 tenors x (1 sub + 4 FMA + 1 add + 1 accumulate)
PF_FLOPS_PER_PATH_STEP = 512       # 51 tenors x 5 FMA + the short-rate sum (portfolio)
PEAK_TFLOPS = {"f32": 157.3, "f64": 78.6}   # MI355X published vector peaks
MFMA64_FLOPS = 2 * 16 * 16 * 4     # one v_mfma_f64_16x16x4f64
PATHS, EXPIRY, DT, SEED = 131072, 10.0, 0.01, 7
TENORS_N = 51
KERNEL_NAMES = {"matrix": "ref_paths_mfma_kernel", "vector": "ref_paths_kernel<float,compensated,philox>"}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def run_plan(hjm, market, swap, precision, steps, warmup, path_kernel=0):
    plan = hjm.ReferencePlan(market, swap, PATHS, DT, precision=precision, seed=SEED,
                             path_kernel=path_kernel)
    for _ in range(warmup):
        plan.execute()
    kernel_ms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        plan.execute()
        kernel_ms.append(plan.timings_ms.copy())
    wall = time.perf_counter() - t0
    out = dict(ee=plan.ee.copy(), pfe=plan.pfe.copy(), cva=float(plan.cva[0]),
               se=float(plan.se[0]), wall=wall, kernel_ms=np.array(kernel_ms))
    plan.close()
    return out


def adjoint_ratio(hjm, market, reps=5):
    """CVA + 204 Greeks versus forward-only device time: adjoint configuration (10y payer,
    65536 paths x 1000 steps, FP32 forward, FP64 accumulation of the Greeks). The
    denominator is a forward-only plan of the same size: path kernel (no stored normals),
    exposure and reductions, no PFE. The Greeks call's own forward also keeps the normals
    for the reverse sweep; both times are reported."""
    swap = hjm.InterestRateSwap(1e6, 0.025, 10.0)
    P = 65536
    hjm.cva_greeks(market, swap, P, DT, precision=hjm.FP32_KAHAN, seed=SEED)
    t = np.array([hjm.cva_greeks(market, swap, P, DT, precision=hjm.FP32_KAHAN,
                                 seed=SEED).timings_ms for _ in range(reps)])
    fwd, adj = float(np.median(t[:, 0])), float(np.median(t[:, 1]))
    plan = hjm.ReferencePlan(market, swap, P, DT, precision=hjm.FP32_KAHAN, seed=SEED,
                             pfe_quantile=0.0)
    plan.execute()
    only = float(np.median([plan.execute().timings_ms[4] for _ in range(reps)]))
    plan.close()
    return {"paths": P, "steps_per_path": 1000, "forward_only_ms": only,
            "greeks_forward_ms": fwd, "adjoint_ms": adj,
            "cva_plus_greeks_over_forward": (fwd + adj) / only,
            "cva_plus_greeks_over_own_forward": (fwd + adj) / fwd, "greeks": 204}


def _eager_timings(hjm, market, pf, paths, tuning, mk, reps=3):
    plan = hjm.PortfolioPlan(market, pf, paths, 30.0, DT, tuning=tuning, **mk)
    plan.execute()
    times = np.array([plan.execute().timings_ms.copy() for _ in range(reps)])
    plan.close()
    return times


def run_portfolio(hjm, args, nswaps, nsets, paths):
    """Portfolio workload: the matrix-core FP32 path kernel, valuation, statistics and PFE
    of every netting set at the launch shapes the autotuner picks for this problem
    (hjm_portfolio_autotune, run before the timed region), replayed as one hipGraph per
    step (timed on the wall clock); the per-kernel breakdown comes from eager runs of the
    same tuning, beside the eager total at the default shapes. After the timed region: the
    same book in FP64 on the same normals (`fp64`: per-pass times with the vector and the
    matrix-core path kernel, and the FP32 run's EE / PFE / CVA errors against FP64)."""
    market = hjm.benchmark_market()
    pf = hjm.synthetic_portfolio(nswaps, nsets, seed=2024)
    S = 3000
    mk = dict(precision=hjm.FP32_KAHAN, seed=2024)
    t0 = time.perf_counter()
    tune, _ = hjm.autotune_portfolio(market, pf, paths, 30.0, DT, precision=hjm.FP32_KAHAN,
                                     seed=2024)
    tune_s = time.perf_counter() - t0
    log("[bench] autotuned shapes: %s (%.1f s)" % (tune.as_dict(), tune_s))
    default = _eager_timings(hjm, market, pf, paths, hjm.Tuning(), mk)
    eager_t = hjm.Tuning(**{**tune.as_dict(), "use_graph": False})
    times = _eager_timings(hjm, market, pf, paths, eager_t, mk)
    plan = hjm.PortfolioPlan(market, pf, paths, 30.0, DT, tuning=tune, **mk)
    for _ in range(max(1, args.warmup)):
        plan.execute()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        plan.execute()
    wall = time.perf_counter() - t0
    r32 = plan.execute()
    plan.close()
    kp = float(np.mean(times[:, 0]))
    fp64 = None
    if not args.no_fp64:
        # the same book in FP64 on the same normals: the vector and the matrix-core path
        # kernel (eager, median of 3), and the FP32 run's errors against it
        log("[bench] FP64 legs (vector and matrix-core path kernels)")
        fp64 = {}
        r64 = None
        for name, pk in (("vector", hjm.KERNEL_VECTOR), ("matrix", hjm.KERNEL_MATRIX)):
            p64 = hjm.PortfolioPlan(market, pf, paths, 30.0, DT, precision=hjm.FP64, seed=2024,
                                    path_kernel=pk)
            p64.execute()
            runs = [p64.execute() for _ in range(3)]
            p64.close()
            t = np.median(np.array([r.timings_ms for r in runs]), axis=0)
            fp64[name] = {"paths": float(t[0]), "valuation": float(t[1]),
                          "statistics": float(t[2]), "pfe_select": float(t[3]),
                          "device_total": float(t[4])}
            if r64 is None:
                r64 = runs[-1]
        # (dates whose EE is above 1e-3 of the set's peak; PFE where it is non-zero)
        peak = r64.expected_exposure.max(axis=1, keepdims=True)
        big = r64.expected_exposure > 1e-3 * peak
        pk = big & (r64.pfe > 0)
        fp64["fp32_vs_fp64_error"] = {
            "ee_max_rel": float(np.max(np.abs(r32.expected_exposure[big] / r64.expected_exposure[big] - 1))),
            "ee_max_err_of_set_peak": float(np.max(np.abs(r32.expected_exposure - r64.expected_exposure) / peak)),
            "pfe_max_rel": float(np.max(np.abs(r32.pfe[pk] / r64.pfe[pk] - 1))),
            "cva_max_rel": float(np.max(np.abs(r32.cva / r64.cva - 1))),
            "cva_se_rel_min": float(np.min(r64.cva_stderr / r64.cva))}
    adj = None
    if args.config == "full":
        # CVA + 204 Greeks over the whole book (forward keeping its shocks, one-pass reverse
        # sweep) against a forward-only plan of the same size (no PFE)
        log("[bench] full configuration: CVA + 204 Greeks")
        g = hjm.portfolio_cva_greeks(market, pf, paths, 30.0, DT, tuning=eager_t, **mk)
        fo = hjm.PortfolioPlan(market, pf, paths, 30.0, DT, pfe_quantile=0.0, tuning=eager_t,
                               **mk)
        fo.execute()
        only = float(np.median([fo.execute().timings_ms[4] for _ in range(3)]))
        fo.close()
        adj = {"paths": paths, "steps_per_path": S, "forward_only_ms": only,
               "greeks_forward_ms": float(g.timings_ms[0]), "adjoint_ms": float(g.timings_ms[1]),
               "cva_plus_greeks_over_forward": float(g.timings_ms[2]) / only,
               "stored_normals_bytes": g.stored_normals_bytes, "greeks": 204}
    flops = PF_FLOPS_PER_PATH_STEP * paths * S
    achieved = flops / (kp * 1e-3) / 1e12
    value = paths * S * args.steps / wall
    return {
        "metric": "path-steps/s (51 tenors), portfolio exposure + CVA per netting set",
        "value": value, "unit": "path-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic portfolio (seed 2024), in-kernel Philox normals",
        "config": {"workload": "portfolio: %d swaps in %d netting sets, %d paths x %d steps "
                               "(30y, dt=0.01), EE/PFE97.5/CVA per set" % (nswaps, nsets, paths, S),
                   "paths": paths, "steps_per_path": S, "swaps": nswaps, "netting_sets": nsets,
                   "hip_graph": True, "tuning": tune.as_dict(), "autotune_s": tune_s},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK_TFLOPS["f32"],
                     "unit": "TFLOP/s", "frac": achieved / PEAK_TFLOPS["f32"], "traffic": None,
                     "kernel": "pf_paths_mfma_kernel", "kernel_ms": kp,
                     "flops_per_launch": flops},
        # fused valuation-statistics pass (if tuned so): [1] is the valuation with the
        # reductions, [2] the column sums
        "kernel_ms_breakdown": {"paths": kp, "valuation": float(np.mean(times[:, 1])),
                                "statistics": float(np.mean(times[:, 2])),
                                "pfe_select": float(np.mean(times[:, 3])),
                                "device_total_eager": float(np.mean(times[:, 4])),
                                "device_total_eager_default_shapes": float(np.mean(default[:, 4])),
                                "paths_default_shapes": float(np.mean(default[:, 0])),
                                "valuation_default_shapes": float(np.mean(default[:, 1]))},
        "fp64": fp64,
        "adjoint": adj,
        "cpu_baseline": None,
    }


def run_calibration(hjm, args, ncurves=2600, window=500):
    """Rolling-window PCA calibration: every window of 500 daily curves (step 1) of a
    2600-day synthetic history (rank-3 benchmark-loading differences plus noise)."""
    market = hjm.benchmark_market()
    sigma = market.volatilities.reshape(3, TENORS_N)
    curves = hjm.synthetic_history(sigma, ncurves, seed=21, noise=2e-5, base=market.spot_rates)
    nw = ncurves - window + 1
    deg = (0, 2, 3)
    for _ in range(args.warmup):
        hjm.calibrate_pca(curves, window=window, step=1, nwindows=nw, fit_degrees=deg)
    t0 = time.perf_counter()
    ks = []
    for _ in range(args.steps):
        ks.append(hjm.calibrate_pca(curves, window=window, step=1, nwindows=nw,
                                    fit_degrees=deg).timings_ms[0])
    wall = time.perf_counter() - t0
    kms = float(np.mean(ks))
    # CPU baseline: the numpy oracle (LAPACK eigh), one window at a time, ~10 s sample
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import calibration as cal
    done, t_cpu = 0, 0.0
    while t_cpu < 10.0 and done < nw:
        t1 = time.perf_counter()
        cal.pca_window(curves[done:done + window])
        t_cpu += time.perf_counter() - t1
        done += 1
    return {
        "metric": "calibration windows/s (rolling PCA of 500 daily curves, 51 tenors)",
        "value": nw * args.steps / wall, "unit": "windows/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic history (rank-3 benchmark loadings + noise, seed 21)",
        "config": {"workload": "calibration: %d windows x %d curves, step 1, fits (0, 2, 3)"
                               % (nw, window), "windows": nw, "window": window},
        "kernel_ms": kms, "kernel_windows_per_s": nw / (kms * 1e-3),
        "cpu_baseline": {"value": done / t_cpu, "unit": "windows/s", "cores": 1,
                         "kind": "port",
                         "sample": "%d windows, numpy cov + LAPACK eigh per window" % done},
    }


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE
    passes (copy-calibrated), or None when no measurement of that kernel is recorded."""
    path = os.path.join(ROOT, "profiles", "ref_paths_mfma_hbm_traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f)
    except (OSError, ValueError):
        return None
    return rec.get("traffic_bytes") if rec.get("kernel") == kernel else None


def host_cpus():
    """The host's CPUs as this process sees them: model string, nproc, the CPUs in the
    scheduler affinity mask, the cgroup CPU quota (cpu.max) and OMP_NUM_THREADS."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, period = f.read().split()[:2]
            quota = None if q == "max" else int(q) / int(period)
    except (OSError, ValueError):
        pass
    omp = os.environ.get("OMP_NUM_THREADS")
    return {"cpu_model": model, "nproc": os.cpu_count(),
            "cpus_in_affinity": len(os.sched_getaffinity(0)), "cgroup_cpu_quota": quota,
            "omp_num_threads": int(omp) if omp and omp.isdigit() else None}


def all_core_threads(info):
    """Threads for the all-cores baseline: every CPU in the affinity mask, capped by the
    job's CPU share where the host states one (cgroup quota, OMP_NUM_THREADS)."""
    n = info["cpus_in_affinity"]
    if info["cgroup_cpu_quota"]:
        n = min(n, max(1, int(info["cgroup_cpu_quota"])))
    if info["omp_num_threads"]:
        n = min(n, info["omp_num_threads"])
    return max(1, n)


def cpu_baseline(market, target_s=10.0, nthreads=1, info=None):
    """The oracle's reference loop (simulation_cpu.h:413-436) in FP64 on `nthreads` threads
    (OpenMP over paths), timed around the path loop only: the first 1024 x nthreads paths
    of the workload, run repeatedly until ~target_s of timed work."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    orc = oracle_lib.load()
    S = int(round(EXPIRY / DT))
    chunk = min(PATHS, 1024 * nthreads)
    z = orc.normals(SEED, chunk, S)
    reps, ms = 0, 0.0
    while ms < target_s * 1e3 and reps < 64:
        t, _, _ = orc.time_reference_paths(market.spot_rates, market.drifts,
                                           market.volatilities, EXPIRY, 0.5, DT, z,
                                           nthreads=nthreads)
        ms += t
        reps += 1
    out = dict(value=reps * chunk * S / (ms * 1e-3), unit="path-steps/s", cores=nthreads,
               kind="port",
               sample="the first %d of the %d paths x %d steps, timed %d times, FP64 reference "
                      "loop, %d thread(s), normals generated outside the timed region"
                      % (chunk, PATHS, S, reps, nthreads))
    if info:
        out.update(info)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-fp64", action="store_true",
                    help="batched / full: skip the FP64 legs of the book")
    ap.add_argument("--config", default="precision",
                    choices=["precision", "batched", "full", "calibration"])
    ap.add_argument("--no-baseline-form", action="store_true",
                    help="skip the reference-design (plain FP64) GPU run")
    ap.add_argument("--path-kernel", default="matrix", choices=["matrix", "vector"],
                    help="FP32 path kernel timed as the headline: matrix cores or VALU")
    ap.add_argument("--headline-only", action="store_true",
                    help="time only the headline kernel sequence (for counter runs)")
    ap.add_argument("--headline-precision", default="fp32", choices=["fp32", "fp64"],
                    help="with --headline-only: the precision of that sequence (fp64 "
                         "profiles the FP64 path kernels, --path-kernel choosing which)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="run only a plain device copy of MIB MiB (calibrates HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the path shards across GPUs by contiguous path "
              "ranges of one seed; runs with --gpus > 1 are not provided.",
              file=sys.stderr)
        sys.exit(2)

    import cva_hjm_cuda_amd as hjm
    if hjm.device_count() < 1:
        raise SystemExit("no HIP device visible")
    # every library call returns after its device work (or reads back events): torch is
    # only needed for the copy calibration and as an extra device-wide fence
    sync = lambda: None  # noqa: E731
    if not args.headline_only:
        try:
            import torch
            if torch.cuda.is_available():
                sync = torch.cuda.synchronize
        except Exception:
            pass

    if args.copy_calibration:
        # HIP's device-to-device copy kernel moves exactly MIB MiB each way per launch; the
        # HBM counters it produces calibrate those of the path kernels
        import ctypes as C
        hip = C.CDLL("libamdhip64.so")
        nbytes = args.copy_calibration * (1 << 20)
        src, dst = C.c_void_p(), C.c_void_p()
        if hip.hipMalloc(C.byref(src), C.c_size_t(nbytes)) or \
                hip.hipMalloc(C.byref(dst), C.c_size_t(nbytes)):
            raise SystemExit("hipMalloc failed")
        if hip.hipMemset(src, 1, C.c_size_t(nbytes)):
            raise SystemExit("hipMemset failed")
        for _ in range(3):
            if hip.hipMemcpy(dst, src, C.c_size_t(nbytes), 3):   # hipMemcpyDeviceToDevice
                raise SystemExit("hipMemcpy failed")
        hip.hipDeviceSynchronize()
        hip.hipFree(src)
        hip.hipFree(dst)
        print(json.dumps({"copy_calibration_bytes": nbytes, "launches": 3}), flush=True)
        return

    if args.config == "calibration":
        print(json.dumps(run_calibration(hjm, args)), flush=True)
        return
    if args.config != "precision":
        n = {"batched": (256, 32, 262144), "full": (1024, 128, 1048576)}[args.config]
        print(json.dumps(run_portfolio(hjm, args, *n)), flush=True)
        return

    market = hjm.benchmark_market()
    swap = hjm.InterestRateSwap(1e6, 0.025, EXPIRY)
    S = int(round(EXPIRY / DT))
    evolved = S - 50                           # block 0 is frozen at the initial curve

    other = "vector" if args.path_kernel == "matrix" else "matrix"
    kernels = {"matrix": hjm.KERNEL_MATRIX, "vector": hjm.KERNEL_VECTOR}
    if args.headline_only:
        fp64 = args.headline_precision == "fp64"
        f32 = run_plan(hjm, market, swap, hjm.FP64 if fp64 else hjm.FP32_KAHAN, args.steps,
                       args.warmup, kernels[args.path_kernel])
        name = ({"matrix": "ref_paths_mfma64_kernel", "vector": "ref_paths_kernel<double>"}
                if fp64 else KERNEL_NAMES)[args.path_kernel]
        print(json.dumps({"headline_only": True, "kernel": name, "precision": args.headline_precision,
                          "kernel_ms": float(np.mean(f32["kernel_ms"][:, 0])),
                          "path_steps_per_s": PATHS * S * args.steps / f32["wall"]}), flush=True)
        return
    log("[bench] FP64 path kernel")
    f64 = run_plan(hjm, market, swap, hjm.FP64, max(2, args.steps // 2), 1)
    log("[bench] FP64 matrix-core path kernel")
    f64m = run_plan(hjm, market, swap, hjm.FP64, max(2, args.steps // 2), 1, hjm.KERNEL_MATRIX)
    log("[bench] FP32 + compensated, %s path kernel" % other)
    alt = run_plan(hjm, market, swap, hjm.FP32_KAHAN, max(2, args.steps // 2), 1,
                   kernels[other])
    log("[bench] FP32 + compensated, %s path kernel (timed)" % args.path_kernel)
    sync()
    f32 = run_plan(hjm, market, swap, hjm.FP32_KAHAN, args.steps, args.warmup,
                   kernels[args.path_kernel])
    sync()

    value = PATHS * S * args.steps / f32["wall"]
    f64_rate = PATHS * S * f64["kernel_ms"].shape[0] / f64["wall"]
    k32 = float(np.mean(f32["kernel_ms"][:, 0]))
    k64 = float(np.mean(f64["kernel_ms"][:, 0]))
    k64m = float(np.mean(f64m["kernel_ms"][:, 0]))
    f64m_rate = PATHS * S * f64m["kernel_ms"].shape[0] / f64m["wall"]
    kalt = float(np.mean(alt["kernel_ms"][:, 0]))
    flops = FLOPS_PER_PATH_STEP * PATHS * evolved
    achieved = flops / (k32 * 1e-3) / 1e12

    plain = None
    if not args.no_baseline_form:
        log("[bench] reference GPU design, FP64 (plain FP64 GPU form)")
        r = hjm.simulate_reference(market, swap, PATHS, DT, seed=SEED, blockwise=True,
                                   pfe_quantile=0.0)
        t = r.timings_ms
        plain = dict(path_steps_per_s=PATHS * S / (t[4] * 1e-3),
                     kernel_path_steps_per_s=PATHS * S / ((t[0] + t[1]) * 1e-3),
                     device_ms=float(t[4]),
                     cva_rel_err_vs_fp64=abs(r.cva / f64["cva"] - 1.0))

    big = f64["ee"] > 1e-3 * f64["ee"].max()

    def error_vs_fp64(r):
        """EE / PFE97.5 (dates with EE above 1e-3 of its peak) and CVA of an FP32 run
        against the FP64 run on the same in-kernel normals."""
        return dict(
            ee_max_rel=float(np.max(np.abs(r["ee"][big] / f64["ee"][big] - 1))),
            pfe_max_rel=float(np.max(np.abs(r["pfe"][big] / f64["pfe"][big] - 1))),
            cva_rel=abs(r["cva"] / f64["cva"] - 1.0),
            cva_se_rel=f64["se"] / f64["cva"])

    err = error_vs_fp64(f32)

    log("[bench] adjoint configuration: CVA + 204 Greeks vs forward")
    adj = adjoint_ratio(hjm, market)
    cpu = cpu_all = None
    if not args.no_cpu:
        log("[bench] CPU oracle, 1 thread and all cores")
        info = host_cpus()
        cpu = cpu_baseline(market, info=info)
        cpu_all = cpu_baseline(market, nthreads=all_core_threads(info), info=info)
    res = {
        "metric": "path-steps/s (51 tenors) FP32 vs FP64; CVA+204 Greeks time vs forward-only",
        "value": value,
        "unit": "path-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": f32["wall"] * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (Nelson-Siegel curve, parametric vol factors, in-kernel Philox "
                "normals seed 7)",
        "config": {"workload": "precision: 1 payer IRS 10y, %d paths x %d steps (dt=0.01), "
                               "51 tenors, 3 factors, FP32 evolution + Kahan-compensated "
                               "integrals, EE/PFE97.5/CVA in FP64" % (PATHS, S),
                   "paths": PATHS, "steps_per_path": S, "tenors": 51, "factors": 3,
                   "parallelism": "single GPU, one path per lane"},
        "roofline": {"bound": "mfma" if args.path_kernel == "matrix" else "valu",
                     "achieved": achieved, "peak": PEAK_TFLOPS["f32"],
                     "unit": "TFLOP/s", "frac": achieved / PEAK_TFLOPS["f32"],
                     "traffic": measured_traffic(KERNEL_NAMES[args.path_kernel]),
                     "kernel": KERNEL_NAMES[args.path_kernel],
                     "kernel_ms": k32,
                     "flops_per_launch": flops},
        "fp32_" + other: {"kernel": KERNEL_NAMES[other], "kernel_ms": kalt,
                          "path_steps_per_s": PATHS * S * alt["kernel_ms"].shape[0] / alt["wall"],
                          "roofline_frac": flops / (kalt * 1e-3) / 1e12 / PEAK_TFLOPS["f32"],
                          "fp32_vs_fp64_error": error_vs_fp64(alt),
                          "cva_rel_vs_headline": abs(alt["cva"] / f32["cva"] - 1.0)},
        "cpu_baseline": cpu,
        "cpu_baseline_all_cores": cpu_all,
        "fp64": {"path_steps_per_s": f64_rate, "kernel_ms": k64,
                 "roofline_frac": flops / (k64 * 1e-3) / 1e12 / PEAK_TFLOPS["f64"],
                 "fp32_over_fp64": value / f64_rate},
        "fp64_matrix": {"kernel": "ref_paths_mfma64_kernel", "kernel_ms": k64m,
                        "path_steps_per_s": f64m_rate,
                        "roofline_frac": flops / (k64m * 1e-3) / 1e12 / PEAK_TFLOPS["f64"],
                        # the products the kernel issues per 16 paths and evolved block:
                        # 52 v_mfma_f64_16x16x4f64 on each of 5 live M-tiles (20 dates)
                        # less the 15 whose block-map entries are structurally zero
                        "mfma_issued_frac": MFMA64_FLOPS * (52 * 5 - 15) * (PATHS // 16) * (evolved // 50)
                        / (k64m * 1e-3) / 1e12 / PEAK_TFLOPS["f64"],
                        "vs_fp64_vector": {
                            "ee_max_rel": float(np.max(np.abs(f64m["ee"][big] / f64["ee"][big] - 1))),
                            "pfe_max_rel": float(np.max(np.abs(f64m["pfe"][big] / f64["pfe"][big] - 1))),
                            "cva_rel": abs(f64m["cva"] / f64["cva"] - 1.0)},
                        "fp32_over_fp64_matrix": value / f64m_rate},
        "plain_fp64_gpu_form": plain,
        "fp32_vs_fp64_error": err,
        "adjoint": adj,
        "kernel_ms_breakdown": {"paths": k32,
                                "exposure": float(np.mean(f32["kernel_ms"][:, 1])),
                                "reduce": float(np.mean(f32["kernel_ms"][:, 2])),
                                "pfe_select": float(np.mean(f32["kernel_ms"][:, 3])),
                                "device_total": float(np.mean(f32["kernel_ms"][:, 4]))},
    }
    if plain:
        res["fp32_over_plain_fp64_form"] = value / plain["path_steps_per_s"]
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
