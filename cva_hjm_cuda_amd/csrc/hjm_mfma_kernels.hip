// hjm_mfma_kernels.hip -- FP32 path kernels on the matrix cores.
//
// Each 50-step block is the affine map [f_out; integral rows] = W [f_in; phi; 1]
// (hjm_mfma_plan.h). A wave owns 32 paths = 2 N-tiles of 16; per block it issues
// MT x 2 x 55 v_mfma_f32_16x16x4f32 (A = W fragments from LDS, B = the previous block's
// accumulators or freshly generated shocks) while the VALU generates the next shocks. The
// accumulator layout of a 16x16 tile -- lane l holds rows 4 (l / 16) + r, r = 0..3, of
// column l % 16 -- is used directly as the next block's B operand (W's columns are
// permuted to match), so the curve never leaves registers. One workgroup of 8 waves
// shares W in LDS.
//   ref_paths_mfma_kernel: reference semantics (7 M-tiles): the block integrals dA are
//     Kahan-added into (A, C); numeraire slot b+1 is recorded after block b.
//   pf_paths_mfma_kernel: portfolio semantics (4 M-tiles): the curve is checkpointed at
//     every date and the short rate's left-point sum (row 51) extends the log bank account.
#include "hjm_common.h"
#include "hjm_kernels.h"
#include "hjm_mfma_plan.h"
#include "hjm_rng.h"

namespace hjm {

typedef float f4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f4_t mfma16(float a, float b, f4_t c)
{
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

constexpr int kNT = 2;                       // N-tiles (16 paths each) per wave
constexpr int kMfmaBlock = 512;              // 8 waves
constexpr int kMfmaPathsPerBlock = kMfmaBlock / kWave * kNT * 16;

// One 50-step block: acc = W x [fin; shocks of steps b*50 .. b*50+49; 1]. fin is consumed
// (rotated in place). The RNG of step quad q+1 runs on the VALU while the matrix cores
// take quad q's shocks and one chunk of the previous curve; scheduling groups interleave
// the two instruction streams within each wave.
template <int MT>
__device__ __forceinline__ void mfma_block(const float* __restrict__ lds_w, float (&fin)[kNT][16],
                                           f4_t (&acc)[MT][kNT], const MfmaGrid& g,
                                           const int64_t (&path)[kNT], int b, int grp, int lane)
{
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int n = 0; n < kNT; ++n) acc[m][n] = f4_t{0.f, 0.f, 0.f, 0.f};
    // shocks: step 4 q + grp of this block, three factors per step. Lanes past the
    // block's last step hold 0, except the constant 1 of the drift.
    float zc[kNT][3], zn[kNT][3];
    auto gen = [&](int q, float (&z)[kNT][3]) {
        const int t = 4 * q + grp;
        const int tc = t < kMfmaStride ? t : kMfmaStride - 1;
#pragma unroll
        for (int n = 0; n < kNT; ++n) {
            const Normals3 v = step_normals(g.seed_lo, g.seed_hi, (uint32_t)(g.path0 + path[n]),
                                            (uint32_t)(b * kMfmaStride + tc));
            const bool in = t < kMfmaStride;
            z[n][0] = in ? v.z0 : (t == kMfmaStride ? 1.f : 0.f);
            z[n][1] = in ? v.z1 : 0.f;
            z[n][2] = in ? v.z2 : 0.f;
        }
    };
    auto f_chunk = [&](int kc, int reg) {
#pragma unroll
        for (int m = 0; m < MT; ++m) {
            const float a = lds_w[(m * kMfmaKChunks + kc) * kWave + lane];
#pragma unroll
            for (int n = 0; n < kNT; ++n) acc[m][n] = mfma16(a, fin[n][reg], acc[m][n]);
        }
    };
    constexpr int kHead = kMfmaFChunks - kMfmaQuads;  // 3 curve chunks before the loop
    gen(0, zc);
#pragma unroll
    for (int kc = 0; kc < kHead; ++kc) f_chunk(kc, kc);
#pragma unroll
    for (int n = 0; n < kNT; ++n)
#pragma unroll
        for (int i = 0; i < kMfmaQuads; ++i) fin[n][i] = fin[n][i + kHead];
#pragma unroll
    for (int i = 0; i < kHead * MT * kNT; ++i) {
        __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
        __builtin_amdgcn_sched_group_barrier(0x002, MT == 7 ? 9 : 16, 0);
    }
    for (int q = 0; q < kMfmaQuads; ++q) {
        gen(q + 1 < kMfmaQuads ? q + 1 : q, zn);
        f_chunk(kHead + q, 0);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const int kc = kMfmaFChunks + 3 * q + k;
#pragma unroll
            for (int m = 0; m < MT; ++m) {
                const float a = lds_w[(m * kMfmaKChunks + kc) * kWave + lane];
#pragma unroll
                for (int n = 0; n < kNT; ++n) acc[m][n] = mfma16(a, zc[n][k], acc[m][n]);
            }
        }
#pragma unroll
        for (int n = 0; n < kNT; ++n) {
#pragma unroll
            for (int i = 0; i + 1 < kMfmaQuads; ++i) fin[n][i] = fin[n][i + 1];
#pragma unroll
            for (int k = 0; k < 3; ++k) zc[n][k] = zn[n][k];
        }
        // one MFMA, then a share of the VALU work, 8 MT times
#pragma unroll
        for (int i = 0; i < 8 * MT; ++i) {
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
            __builtin_amdgcn_sched_group_barrier(0x002, MT == 7 ? 7 : 12, 0);
        }
    }
}

__global__ __launch_bounds__(kMfmaBlock) void ref_paths_mfma_kernel(const MfmaGrid g,
                                                                const float* __restrict__ wfrag,
                                                                double* __restrict__ num_x,
                                                                double* __restrict__ num_y)
{
    extern __shared__ float lds_w[];
    for (int t = threadIdx.x; t < kMfmaFragFloats; t += kMfmaBlock) lds_w[t] = wfrag[t];
    __syncthreads();

    const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    const int grp = lane >> 4, col = lane & 15;
    const int64_t P = g.paths;
    const int64_t base = (int64_t)blockIdx.x * kMfmaPathsPerBlock + (int64_t)wv * kNT * 16;
    int64_t path[kNT];
    bool live[kNT];
#pragma unroll
    for (int n = 0; n < kNT; ++n) {
        path[n] = base + 16 * n + col;
        live[n] = path[n] < P;
    }

    // curve rows 16 m + 4 grp + r of the N-tiles (B operand of the F chunks)
    float fin[kNT][16];
    // Kahan (A, C) of the rows of tiles 3..6 (dA_j at row 51 + j)
    float Ak[kNT][16], Ck[kNT][16];
#pragma unroll
    for (int n = 0; n < kNT; ++n) {
#pragma unroll
        for (int m = 0; m < 4; ++m)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int R = 16 * m + 4 * grp + r;
                fin[n][4 * m + r] = (R < kTenors) ? g.spot[R < kTenors ? R : 0] : 0.f;
            }
#pragma unroll
        for (int m = 3; m < 7; ++m)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int j = 16 * m + 4 * grp + r - kTenors;
                const double a0 = (j >= 0 && j < kTenors) ? g.a0[j >= 0 && j < kTenors ? j : 0] : 0.0;
                const float af = (float)a0;
                Ak[n][4 * (m - 3) + r] = af;
                Ck[n][4 * (m - 3) + r] = (float)((double)af - a0);
            }
    }
    if (grp == 0) {
#pragma unroll
        for (int n = 0; n < kNT; ++n)
            if (live[n]) {
                num_x[P + path[n]] = (double)g.spot[1];
                num_y[P + path[n]] = exp(-g.a0[1] * g.dt);
            }
    }

    for (int b = 1; b < g.ndates; ++b) {
        f4_t acc[kMfmaMTiles][kNT];
        mfma_block<kMfmaMTiles>(lds_w, fin, acc, g, path, b, grp, lane);
        // integrals: Kahan-add the block's dA (tiles 3..6); curve rows become next input
#pragma unroll
        for (int n = 0; n < kNT; ++n) {
#pragma unroll
            for (int m = 3; m < 7; ++m)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int i = 4 * (m - 3) + r;
                    const float y = acc[m][n][r] - Ck[n][i];
                    const float tt = Ak[n][i] + y;
                    Ck[n][i] = (tt - Ak[n][i]) - y;
                    Ak[n][i] = tt;
                }
#pragma unroll
            for (int m = 0; m < 4; ++m)
#pragma unroll
                for (int r = 0; r < 4; ++r) fin[n][4 * m + r] = acc[m][n][r];
        }
        // slot b+1: x = curve row b+1, integral = row 51 + b + 1 (uniform positions)
        const int slot = b + 1;
        const int Rx = slot, Ra = kTenors + slot;
        const int gx = (Rx % 16) / 4, ix = 4 * (Rx / 16) + Rx % 4;
        const int ga = (Ra % 16) / 4, ia = 4 * (Ra / 16 - 3) + Ra % 4;
#pragma unroll
        for (int n = 0; n < kNT; ++n) {
            float xv = 0.f, av = 0.f, cv = 0.f;
#pragma unroll
            for (int t2 = 0; t2 < 16; ++t2) {
                if (t2 == ix) xv = fin[n][t2];
                if (t2 == ia) {
                    av = Ak[n][t2];
                    cv = Ck[n][t2];
                }
            }
            if (live[n]) {
                if (grp == gx) num_x[(int64_t)slot * P + path[n]] = (double)xv;
                if (grp == ga)
                    num_y[(int64_t)slot * P + path[n]] = exp(-((double)av - (double)cv) * g.dt);
            }
        }
    }
}

// Portfolio semantics: dates e = 0 .. ndates-1 every 50 steps from t = 0. At each date the
// curve goes to curves[e][tenor][path] and the log bank account dt sum_{n < 50 e} f_n(0)
// to logb[e][path]; the block's row 51 (left-point sum of the short rate) extends it.
__global__ __launch_bounds__(kMfmaBlock) void pf_paths_mfma_kernel(const MfmaGrid g,
                                                               const float* __restrict__ wfrag,
                                                               float* __restrict__ curves,
                                                               double* __restrict__ logb)
{
    extern __shared__ float lds_w[];
    for (int t = threadIdx.x; t < kMfmaPfFragFloats; t += kMfmaBlock) lds_w[t] = wfrag[t];
    __syncthreads();

    const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    const int grp = lane >> 4, col = lane & 15;
    const int64_t P = g.paths;
    const int64_t base = (int64_t)blockIdx.x * kMfmaPathsPerBlock + (int64_t)wv * kNT * 16;
    int64_t path[kNT];
    bool live[kNT];
    float fin[kNT][16];
    double lnB[kNT];
#pragma unroll
    for (int n = 0; n < kNT; ++n) {
        path[n] = base + 16 * n + col;
        live[n] = path[n] < P;
        lnB[n] = 0.0;
#pragma unroll
        for (int m = 0; m < 4; ++m)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int R = 16 * m + 4 * grp + r;
                fin[n][4 * m + r] = (R < kTenors) ? g.spot[R < kTenors ? R : 0] : 0.f;
            }
    }
    // row 51 (the short-rate sum) is register 4 * 3 + 3 of lane group 0
    constexpr int kSumReg = 4 * (kTenors / 16) + kTenors % 4;
    const bool sum_lane = grp == (kTenors % 16) / 4;
    for (int e = 0;; ++e) {
#pragma unroll
        for (int n = 0; n < kNT; ++n) {
            if (!live[n]) continue;
#pragma unroll
            for (int m = 0; m < 4; ++m)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int R = 16 * m + 4 * grp + r;
                    if (R < kTenors) curves[((int64_t)e * kTenors + R) * P + path[n]] = fin[n][4 * m + r];
                }
            if (sum_lane) logb[(int64_t)e * P + path[n]] = lnB[n];
        }
        if (e + 1 >= g.ndates) break;
        f4_t acc[kMfmaPfMTiles][kNT];
        mfma_block<kMfmaPfMTiles>(lds_w, fin, acc, g, path, e, grp, lane);
#pragma unroll
        for (int n = 0; n < kNT; ++n) {
#pragma unroll
            for (int m = 0; m < 4; ++m)
#pragma unroll
                for (int r = 0; r < 4; ++r) fin[n][4 * m + r] = acc[m][n][r];
            lnB[n] += (double)fin[n][kSumReg] * g.dt;
        }
    }
}

hipError_t mfma_kernels_prepare()
{
    hipError_t e = hipFuncSetAttribute((const void*)ref_paths_mfma_kernel,
                                       hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)(sizeof(float) * kMfmaFragFloats));
    if (e != hipSuccess) return e;
    return hipFuncSetAttribute((const void*)pf_paths_mfma_kernel,
                               hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)(sizeof(float) * kMfmaPfFragFloats));
}

hipError_t launch_ref_paths_mfma(const MfmaGrid& g, const float* wfrag, double* nx, double* ny,
                                 hipStream_t st)
{
    const dim3 grid((unsigned)((g.paths + kMfmaPathsPerBlock - 1) / kMfmaPathsPerBlock));
    ref_paths_mfma_kernel<<<grid, kMfmaBlock, sizeof(float) * kMfmaFragFloats, st>>>(g, wfrag, nx, ny);
    return hipGetLastError();
}

hipError_t launch_pf_paths_mfma(const MfmaGrid& g, const float* wfrag, float* curves, double* logb,
                                hipStream_t st)
{
    const dim3 grid((unsigned)((g.paths + kMfmaPathsPerBlock - 1) / kMfmaPathsPerBlock));
    pf_paths_mfma_kernel<<<grid, kMfmaBlock, sizeof(float) * kMfmaPfFragFloats, st>>>(g, wfrag, curves,
                                                                                      logb);
    return hipGetLastError();
}

}  // namespace hjm
