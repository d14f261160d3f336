// hjm_mfma64_kernels.hip -- the FP64 path kernels on the matrix cores
// (v_mfma_f64_16x16x4f64).
//
// Each 50-step block is the affine map [f_out; integral rows] = W [f_in; phi; 1] of the
// FP32 kernels, held here in FP64 (build_mfma64_block_map) and applied with one FP64 product
// per group of 4 K entries: no operand splitting. The 16-row x 16-path accumulator tile
// (lane l: rows 4 i + l / 16 of path l % 16 in register i) is the next block's B operand:
// register i of curve tile m feeds instruction 4 m + i, so the curve stays in registers.
// The shocks are the FP32 kernels' (same streams, same values), converted exactly to FP64.
// On gfx950 an FP64 product holds the matrix pipe 64 cycles (78.6 TFLOP/s, the FP64 VALU's
// rate) and does not co-execute with VALU work, so the block's products cost what the
// step-by-step kernel's FP64 FMAs do; the gain is utilisation: a wave's M-tiles are
// independent accumulators that keep the pipe busy, where one path per lane waits on a
// dependent chain of steps. A wave owns 16 paths; a workgroup of 8 to 16 waves stages W's
// first kMfma64LdsTiles M-tiles in LDS (26 KB each) and reads the rest, present only for
// reference runs of 29 or more dates, from global memory (L2-resident).
//   ref_paths_mfma64_kernel: reference semantics (up to 7 M-tiles, numeraire slots).
//   pf_paths_mfma64_kernel: portfolio semantics (4 M-tiles, curve checkpoints and the log
//     bank account at every date).
#include "hjm_common.h"
#include "hjm_kernels.h"
#include "hjm_mfma_plan.h"
#include "hjm_rng.h"

#include <utility>

namespace hjm {

typedef double d4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ d4_t mfma_f64(double a, double b, d4_t c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// waves per workgroup (one workgroup per CU: W's tiles fill the LDS): 4 per SIMD where the
// registers allow (up to 5 M-tiles, 116-124 VGPRs), 3 above (152-166); the variants that
// store their shocks at 4 and 7 tiles take one step fewer (they spilled at the larger count)
__host__ __device__ constexpr int m64_waves(int mt, bool store)
{
    return mt <= 5 ? (store && mt == 4 ? 12 : 16) : (store && mt == 7 ? 8 : 12);
}

template <class F, int... Q>
__device__ __forceinline__ void for_each_quad(F& f, std::integer_sequence<int, Q...>)
{
    (f(std::integral_constant<int, Q>{}), ...);
}

// One 50-step block of the wave's 16 paths: acc = W x [fin; shocks of steps b*50 ..
// b*50+49; 1]. kStore: also write the block's normals to zs[step][factor][path] (FP32).
template <int MT, bool kStore>
__device__ __forceinline__ void mfma64_block(const double* __restrict__ lds_w,
                                             const double* __restrict__ wglob, const double (&fin)[16],
                                             d4_t (&acc)[MT], const MfmaGrid& g, int64_t path,
                                             bool live, int b, int grp, int lane,
                                             float* __restrict__ zs)
{
    constexpr int kLds = MT < kMfma64LdsTiles ? MT : kMfma64LdsTiles;
#pragma unroll
    for (int m = 0; m < MT; ++m) acc[m] = d4_t{0.0, 0.0, 0.0, 0.0};
    // W is the same in every block: an opaque offset keeps its loads inside the block
    uint32_t off = (uint32_t)lane;
    asm volatile("" : "+v"(off));
    // instruction j on every M-tile (independent accumulators, one per tile) but those
    // whose W entries are structurally zero (j and m are constants once unrolled)
    auto column = [&](int j, double x) {
#pragma unroll
        for (int m = 0; m < MT; ++m) {
            if (mfma64_column_zero(m, j)) continue;
            const double a = m < kLds ? lds_w[(m * kMfma64K + j) * kWave + off]
                                      : wglob[(m * kMfma64K + j) * kWave + off];
            acc[m] = mfma_f64(a, x, acc[m]);
        }
    };
#pragma unroll
    for (int j = 0; j < kMfma64Curve; ++j) column(j, fin[j]);
    // shocks: value v = 3 q + k of lane group grp is factor k of step 4 q + grp; step 50
    // (lane group 2, q = 12) carries the constant 1 of the drift in factor 0
    float z[4 * ((3 * kMfmaQuads + 3) / 4)];
    auto quad = [&](auto qc) {
        constexpr int q = decltype(qc)::value;
        constexpr bool last = 4 * q + 3 >= kMfmaStride;
        const int t = 4 * q + grp;
        const bool in = !last || t < kMfmaStride;
        // the calls this quad needs first (the lane group's class stream, as the FP32
        // kernels draw it)
        constexpr int chi = (3 * q + 2) / 4;
        constexpr int clo = q == 0 ? 0 : (3 * q - 1) / 4 + 1;
#pragma unroll
        for (int c = clo; c <= chi; ++c) {
            float v[4];
            stream_call(g.seed_lo, g.seed_hi, (uint32_t)(g.path0 + path), (uint32_t)b, (uint32_t)grp,
                        (uint32_t)c, v);
#pragma unroll
            for (int e = 0; e < 4; ++e) z[4 * c + e] = v[e];
        }
        if constexpr (last) {
            z[3 * q + 0] = in ? z[3 * q + 0] : (t == kMfmaStride ? 1.f : 0.f);
            z[3 * q + 1] = in ? z[3 * q + 1] : 0.f;
            z[3 * q + 2] = in ? z[3 * q + 2] : 0.f;
        }
        if constexpr (kStore) {
            if (in && live) {
                float* o = zs + (int64_t)(b * kMfmaStride + t) * 3 * g.paths + path;
                o[0] = z[3 * q + 0];
                o[g.paths] = z[3 * q + 1];
                o[2 * g.paths] = z[3 * q + 2];
            }
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) column(kMfma64Curve + 3 * q + k, (double)z[3 * q + k]);
    };
    for_each_quad(quad, std::make_integer_sequence<int, kMfmaQuads>{});
}

// Reference semantics, as ref_paths_mfma_kernel: MT = the live M-tiles (ref_mfma_tiles);
// block 0 keeps the initial curve (its integrals a0 come from the host), block b >= 1
// evolves 50 steps and records numeraire slot b + 1; the integrals are plain FP64 sums.
template <bool kStore, int MT>
__global__ __launch_bounds__(m64_waves(MT, kStore) * kWave) void ref_paths_mfma64_kernel(const MfmaGrid g,
                                                                  const double* __restrict__ wfrag,
                                                                  double* __restrict__ num_x,
                                                                  double* __restrict__ num_y,
                                                                  float* __restrict__ zs)
{
    constexpr int kLds = MT < kMfma64LdsTiles ? MT : kMfma64LdsTiles;
    constexpr int kInt = MT - 3;  // M-tiles holding integral rows (tile 3 holds both kinds)
    constexpr int kThreads = m64_waves(MT, kStore) * kWave;
    extern __shared__ double lds_w64[];
    for (int t = threadIdx.x; t < kLds * kMfma64K * kWave; t += kThreads) lds_w64[t] = wfrag[t];
    __syncthreads();

    const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    const int grp = lane >> 4, col = lane & 15;
    const int64_t P = g.paths;
    const int64_t path = (int64_t)blockIdx.x * (kThreads / 4) + (int64_t)wv * 16 + col;
    const bool live = path < P;

    // register r of tile m: row 16 m + 4 r + grp -- curve rows (tiles 0..3), integrals of
    // rows 51 + j (tiles 3 .. MT-1)
    double fin[16], A[4 * kInt];
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int R = 16 * m + 4 * r + grp;
            fin[4 * m + r] = R < kTenors ? g.spot64[R < kTenors ? R : 0] : 0.0;
        }
#pragma unroll
    for (int m = 3; m < MT; ++m)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int j = 16 * m + 4 * r + grp - kTenors;
            A[4 * (m - 3) + r] = (j >= 0 && j < kTenors) ? g.a0[j >= 0 && j < kTenors ? j : 0] : 0.0;
        }
    if (grp == 0 && live) {
        num_x[P + path] = g.spot64[1];
        num_y[P + path] = exp(-g.a0[1] * g.dt);
    }

    const double* wglob = wfrag;
    for (int b = 1; b < g.ndates; ++b) {
        d4_t acc[MT];
        mfma64_block<MT, kStore>(lds_w64, wglob, fin, acc, g, path, live, b, grp, lane, zs);
#pragma unroll
        for (int m = 3; m < MT; ++m)
#pragma unroll
            for (int r = 0; r < 4; ++r) A[4 * (m - 3) + r] += acc[m][r];
#pragma unroll
        for (int m = 0; m < 4; ++m)
#pragma unroll
            for (int r = 0; r < 4; ++r) fin[4 * m + r] = acc[m][r];
        // slot b+1: x = curve row b+1, integral = row 51 + b + 1
        const int slot = b + 1;
        const int Rx = slot, Ra = kTenors + slot;
        const int gx = Rx % 4, ix = 4 * (Rx / 16) + (Rx % 16) / 4;
        const int ga = Ra % 4, ia = 4 * (Ra / 16 - 3) + (Ra % 16) / 4;
        double xv = 0.0, av = 0.0;
#pragma unroll
        for (int t2 = 0; t2 < 16; ++t2)
            if (t2 == ix) xv = fin[t2];
#pragma unroll
        for (int t2 = 0; t2 < 4 * kInt; ++t2)
            if (t2 == ia) av = A[t2];
        if (live) {
            if (grp == gx) num_x[(int64_t)slot * P + path] = xv;
            if (grp == ga) num_y[(int64_t)slot * P + path] = exp(-av * g.dt);
        }
    }
}

// Portfolio semantics, as pf_paths_mfma_kernel in FP64: dates e = 0 .. ndates-1 every 50
// steps from t = 0. At each date the curve goes to curves[e][tenor][path] (FP64) and the log
// bank account dt sum_{n < 50 e} f_n(0) to logb[e][path]; block e's row 51 (the short rate's
// left-point sum: register 0 of lane group 3 in tile 3) extends it.
template <bool kStore>
__global__ __launch_bounds__(m64_waves(kMfmaPfMTiles, kStore) * kWave) void pf_paths_mfma64_kernel(
    const MfmaGrid g, const double* __restrict__ wfrag, double* __restrict__ curves,
    double* __restrict__ logb, float* __restrict__ zs)
{
    constexpr int MT = kMfmaPfMTiles;
    constexpr int kThreads = m64_waves(MT, kStore) * kWave;
    extern __shared__ double lds_w64[];
    for (int t = threadIdx.x; t < MT * kMfma64K * kWave; t += kThreads) lds_w64[t] = wfrag[t];
    __syncthreads();

    const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    const int grp = lane >> 4, col = lane & 15;
    const int64_t P = g.paths;
    const int64_t path = (int64_t)blockIdx.x * (kThreads / 4) + (int64_t)wv * 16 + col;
    const bool live = path < P;
    // register r of tile m: row 16 m + 4 r + grp (curve rows below 51; row 51 the sum)
    double fin[16];
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int R = 16 * m + 4 * r + grp;
            fin[4 * m + r] = R < kTenors ? g.spot64[R < kTenors ? R : 0] : 0.0;
        }
    double lnB = 0.0;
    constexpr int kSumReg = 4 * 3 + 0;
    const bool sum_lane = grp == kTenors % 4;
    for (int e = 0;; ++e) {
        if (live) {
#pragma unroll
            for (int m = 0; m < 4; ++m)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int R = 16 * m + 4 * r + grp;
                    if (R < kTenors) curves[((int64_t)e * kTenors + R) * P + path] = fin[4 * m + r];
                }
            if (sum_lane) logb[(int64_t)e * P + path] = lnB;
        }
        if (e + 1 >= g.ndates) break;
        d4_t acc[MT];
        mfma64_block<MT, kStore>(lds_w64, wfrag, fin, acc, g, path, live, e, grp, lane, zs);
#pragma unroll
        for (int m = 0; m < 4; ++m)
#pragma unroll
            for (int r = 0; r < 4; ++r) fin[4 * m + r] = acc[m][r];
        lnB += fin[kSumReg] * g.dt;
    }
}

static size_t mfma64_lds_bytes(int mt)
{
    return (size_t)(mt < kMfma64LdsTiles ? mt : kMfma64LdsTiles) * kMfma64TileBytes;
}

hipError_t mfma64_kernels_prepare()
{
    const void* fns[10] = {
        (const void*)ref_paths_mfma64_kernel<false, 4>, (const void*)ref_paths_mfma64_kernel<false, 5>,
        (const void*)ref_paths_mfma64_kernel<false, 6>, (const void*)ref_paths_mfma64_kernel<false, 7>,
        (const void*)ref_paths_mfma64_kernel<true, 4>,  (const void*)ref_paths_mfma64_kernel<true, 5>,
        (const void*)ref_paths_mfma64_kernel<true, 6>,  (const void*)ref_paths_mfma64_kernel<true, 7>,
        (const void*)pf_paths_mfma64_kernel<false>,      (const void*)pf_paths_mfma64_kernel<true>};
    for (int i = 0; i < 10; ++i) {
        const int mt = i < 8 ? 4 + i % 4 : kMfmaPfMTiles;
        const hipError_t e = hipFuncSetAttribute(fns[i], hipFuncAttributeMaxDynamicSharedMemorySize,
                                                 (int)mfma64_lds_bytes(mt));
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// wfrag: build_mfma64_block_map's fragments of at least ref_mfma_tiles(g.ndates) M-tiles.
hipError_t launch_ref_paths_mfma64(const MfmaGrid& g, const double* wfrag, double* nx, double* ny,
                                   float* zs, hipStream_t st)
{
    if (g.ndates < 1 || g.ndates > kTenors - 1 || g.paths <= 0) return hipErrorInvalidValue;
    const int mt = ref_mfma_tiles(g.ndates);
    const int threads = m64_waves(mt, zs != nullptr) * kWave, per_block = threads / 4;  // 16 paths per wave
    const dim3 grid((unsigned)((g.paths + per_block - 1) / per_block));
    const size_t lds = mfma64_lds_bytes(mt);
#define HJM_REF64_LAUNCH(MT)                                                                   \
    (zs ? ref_paths_mfma64_kernel<true, MT><<<grid, threads, lds, st>>>(g, wfrag, nx, ny, zs)  \
        : ref_paths_mfma64_kernel<false, MT><<<grid, threads, lds, st>>>(g, wfrag, nx, ny, zs))
    switch (mt) {
    case 4: HJM_REF64_LAUNCH(4); break;
    case 5: HJM_REF64_LAUNCH(5); break;
    case 6: HJM_REF64_LAUNCH(6); break;
    default: HJM_REF64_LAUNCH(7); break;
    }
#undef HJM_REF64_LAUNCH
    return hipGetLastError();
}

// wfrag: build_mfma64_block_map's portfolio fragments (4 M-tiles); curves [ndates][51][paths]
// and logb [ndates][paths] in FP64.
hipError_t launch_pf_paths_mfma64(const MfmaGrid& g, const double* wfrag, double* curves, double* logb,
                                  float* zs, hipStream_t st)
{
    if (g.ndates < 1 || g.ndates > kMaxDates || g.paths <= 0) return hipErrorInvalidValue;
    const int threads = m64_waves(kMfmaPfMTiles, zs != nullptr) * kWave, per_block = threads / 4;
    const dim3 grid((unsigned)((g.paths + per_block - 1) / per_block));
    const size_t lds = mfma64_lds_bytes(kMfmaPfMTiles);
    if (zs)
        pf_paths_mfma64_kernel<true><<<grid, threads, lds, st>>>(g, wfrag, curves, logb, zs);
    else
        pf_paths_mfma64_kernel<false><<<grid, threads, lds, st>>>(g, wfrag, curves, logb, zs);
    return hipGetLastError();
}

}  // namespace hjm
