// hjm_common.h -- shared definitions of the HJM exposure library (gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "hjm_cva.h"

namespace hjm {

constexpr int kTenors = HJM_TENORS;    // 51-point Musiela grid, tau_j = j * dtau
constexpr int kFactors = HJM_FACTORS;  // 3 volatility factors
constexpr int kWave = 64;              // CDNA wavefront
constexpr int kBlock = 256;            // one path per lane, 4 waves per workgroup
constexpr int kStatSlots = kTenors + 2;  // EE sums (51), per-path CVA sum and square sum

// c = dt / dtau as the FP32 Euler step uses it: rounded to a multiple of 2^-23, so that
// 1 - c and 1 + c are exact in FP32 and the coefficient pairs (1 - c, c) of the interior
// update f' = (1 - c) f_j + c f_{j+1} + ... and (1 + c, -c) of the last tenor sum to 1
// exactly. With c and 1 - c rounded independently (0.02 and 0.98 sum to 1 + 1.9e-8 in
// FP32) every step scales a flat curve by 1 + 1.9e-8: a systematic +5.8e-7 on the
// forwards after 950 steps and a 2.7e-5 error of the 10y CVA, where the rounded pair
// leaves 6e-8 (CPU emulation of the FP32 step against the FP64 oracle, 2048 paths).
inline float step_ratio_f32(double c) { return (float)(nearbyint(c * 8388608.0) / 8388608.0); }

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
    double spot64[kTenors];  // the FP64 kernel's initial curve
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
    // [date][set] the set's largest live period (0: every swap of the set has matured, its
    // values and exposures at that date are 0), or null
    const int* kmax;
};

// ---- wave / block reductions (fixed order, deterministic) ------------------------
// v + (v moved by the DPP control kCtrl), in the rows of kRowMask (other lanes add 0)
template <int kCtrl, int kRowMask>
__device__ __forceinline__ double dpp_add(double v)
{
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)b, kCtrl, kRowMask, 0xf, false);
    const uint32_t hi =
        (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(b >> 32), kCtrl, kRowMask, 0xf, false);
    return v + __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// Sum over the 64 lanes, returned in every lane: quad butterflies, row shifts by 4 and 8
// (lane 15 of each row of 16 then holds the row's sum), row broadcasts 15 and 31 (lane 63
// holds the wave's sum), read back from lane 63. DPP moves between VALU operations: no
// LDS traffic, unlike ds_bpermute shuffles.
__device__ __forceinline__ double wave_sum(double v)
{
    v = dpp_add<0xb1, 0xf>(v);   // quad_perm [1, 0, 3, 2]
    v = dpp_add<0x4e, 0xf>(v);   // quad_perm [2, 3, 0, 1]
    v = dpp_add<0x114, 0xf>(v);  // row_shr:4
    v = dpp_add<0x118, 0xf>(v);  // row_shr:8
    v = dpp_add<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
    v = dpp_add<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, kWave - 1);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), kWave - 1);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

}  // namespace hjm
