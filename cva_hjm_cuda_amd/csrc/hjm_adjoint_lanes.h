// hjm_adjoint_lanes.h -- device helpers of the one-pass adjoint sweeps, in which two lanes
// share a path: the lower half-wave holds tenors 0..25, the upper half-wave tenors 25..50,
// its tenor 25 a copy taken from the lower lane every step (v_permlane32_swap).
#pragma once
#include "hjm_common.h"
#include "hjm_rng.h"

namespace hjm {

constexpr int kHalfSlots = 26;               // tenor slots per lane (half a curve + 1)
constexpr int kAdjPaths = kBlock / 2;        // paths per workgroup: two lanes per path
constexpr int kAdjKinds = 5;                 // g0, g1, g2, sum of lambda_e, d spot

// FP32 sum over each half-wave (quad butterflies, row shifts by 4 and 8, row broadcast 15
// into rows 1 and 3): lane 31 receives the sum of lanes 0..31, lane 63 that of 32..63.
template <int kCtrl, int kRowMask>
__device__ __forceinline__ float dpp_add_f(float v)
{
    const int m = __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), kCtrl, kRowMask, 0xf, false);
    return v + __builtin_bit_cast(float, m);
}
__device__ __forceinline__ float half_wave_sum(float v)
{
    v = dpp_add_f<0xb1, 0xf>(v);   // quad_perm [1, 0, 3, 2]
    v = dpp_add_f<0x4e, 0xf>(v);   // quad_perm [2, 3, 0, 1]
    v = dpp_add_f<0x114, 0xf>(v);  // row_shr:4
    v = dpp_add_f<0x118, 0xf>(v);  // row_shr:8
    v = dpp_add_f<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
    return v;
}

// The shocks of the steps of a 16-step group, shared by a path's two lanes: the lower lane
// makes the calls of classes 0 and 1, the upper lane those of classes 2 and 3, each into
// the path's LDS column (the layout of fill_group).
__device__ __forceinline__ void fill_group_half(uint32_t seed_lo, uint32_t seed_hi, uint32_t path,
                                                uint32_t block, int m, int half, float* col, int ld)
{
#pragma unroll 1
    for (int g = 2 * half; g < 2 * half + 2; ++g)
#pragma unroll 1
        for (int cc = 0; cc < 3; ++cc) {
            float v[4];
            stream_call(seed_lo, seed_hi, path, block, (uint32_t)g, (uint32_t)(3 * m + cc), v);
#pragma unroll
            for (int e = 0; e < 4; ++e) col[(12 * g + 4 * cc + e) * ld] = v[e];
        }
}

}  // namespace hjm
