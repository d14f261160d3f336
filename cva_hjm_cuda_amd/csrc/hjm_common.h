// hjm_common.h -- shared definitions of the HJM exposure library (gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hjm_cva.h"

namespace hjm {

constexpr int kTenors = HJM_TENORS;    // 51-point Musiela grid, tau_j = j * dtau
constexpr int kFactors = HJM_FACTORS;  // 3 volatility factors
constexpr int kWave = 64;              // CDNA wavefront
constexpr int kBlock = 256;            // one path per lane, 4 waves per workgroup
constexpr int kStatSlots = kTenors + 2;  // EE sums (51), per-path CVA sum and square sum

// Market constants of the reference-semantics path kernel, passed by value so that
// they live in the kernarg segment and reach the ALUs through scalar loads: every lane
// of a wave reads the same tenor constant at the same time (the role of the
// reference's __constant__ arrays, simulation_gpu.cu:109-119).
template <typename Real>
struct RefStepConsts {
    Real ddt[kTenors];   // drift_j * dt
    Real v0[kTenors];    // factor loadings, factor-major (simulation_gpu.cu:429-431)
    Real v1[kTenors];
    Real v2[kTenors];
    Real spot[kTenors];  // f(0, tau_j)
    Real drift[kTenors]; // m(tau_j) (the baseline kernel keeps the reference's op order)
    Real c;              // dt / dtau (upwind dF/dtau * dt coefficient)
    Real omc;            // 1 - c
    Real opc;            // 1 + c
    Real sqrt_dt;
    Real dtau;
};

struct RefGrid {
    int stride;        // Euler steps per exposure block = dtau / dt (50)
    int nsteps;        // expiry / dt
    int ndates;        // nsteps / stride (numeraire slots 1..ndates)
    int pad;
    int64_t paths;     // paths in this launch
    int64_t path0;     // global id of the first path (Philox counter)
    uint32_t seed_lo, seed_hi;
    double dt;
    double a0[kTenors];  // frozen block-0 integral: spot_j summed stride times (FP64)
};

// Matrix-core FP32 path kernel (reference semantics, 50-step blocks).
struct MfmaGrid {
    int ndates;
    int pad;
    int64_t paths;
    int64_t path0;
    uint32_t seed_lo, seed_hi;
    double dt;
    double a0[kTenors];   // frozen block-0 integrals (FP64)
    float spot[kTenors];
};

// Portfolio semantics: dates every dtau up to kMaxPeriods periods (30y at dtau = 0.5).
constexpr int kMaxPeriods = 60;
constexpr int kMaxDates = kMaxPeriods + 1;

struct PfGrid {
    int stride;      // Euler steps per date interval
    int nsteps;      // horizon / dt
    int ndates_m1;   // index of the last date (horizon / dtau)
    int pad;
    int64_t paths;
    int64_t path0;
    uint32_t seed_lo, seed_hi;
    double dt;
};

struct PfValueParams {
    int64_t paths;
    int nsets;
    int ndates;
    double dtau;
};

// ---- wave / block reductions (fixed order, deterministic) ------------------------
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

}  // namespace hjm
