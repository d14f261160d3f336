"""Build libhjm_cva.so (HIP, gfx950) in-tree with hipcc.

Usage: python -m cva_hjm_cuda_amd.build [--verbose]
Output: cva_hjm_cuda_amd/lib/libhjm_cva.so and the command-line driver
cva_hjm_cuda_amd/lib/hjm_cli (host C++ over the library, rpath $ORIGIN).
"""
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "lib")
LIB = os.path.join(LIBDIR, "libhjm_cva.so")
CLI = os.path.join(LIBDIR, "hjm_cli")
SOURCES = ["hjm_capi.hip", "hjm_ref_kernels.hip", "hjm_baseline_kernels.hip",
           "hjm_adjoint_kernels.hip", "hjm_portfolio_kernels.hip",
           "hjm_capi_portfolio.hip", "hjm_mfma_kernels.hip", "hjm_mfma64_kernels.hip", "hjm_calibration.hip",
           "hjm_select_kernels.hip"]
ARCH = os.environ.get("HJM_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared",
    "--offload-arch=" + ARCH,
    # no packed FP32 instructions anywhere in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose=False, extra=()):
    os.makedirs(LIBDIR, exist_ok=True)
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + list(extra) + [
        "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
        "-o", tmp] + [os.path.join(CSRC, s) for s in SOURCES]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed with exit code %d" % res.returncode)
    os.replace(tmp, LIB)
    build_cli(verbose)
    return LIB


def build_cli(verbose=False):
    tmp = CLI + ".tmp"
    cmd = [os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-Wall", "-Wextra",
           "-I" + os.path.join(ROOT, "include"), "-o", tmp, os.path.join(CSRC, "hjm_cli.cpp"),
           "-L" + LIBDIR, "-lhjm_cva", "-Wl,-rpath,$ORIGIN"]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("g++ (driver) failed with exit code %d" % res.returncode)
    os.replace(tmp, CLI)
    return CLI


if __name__ == "__main__":
    print(build(verbose="--verbose" in sys.argv))
