// hjm_mfma_kernels.hip -- FP32 path, valuation and adjoint-seed kernels on the matrix cores.
//
// Each 50-step block is the affine map [f_out; integral rows] = W [f_in; phi; 1]
// (hjm_mfma_plan.h), computed with v_mfma_f32_16x16x32_bf16 as six products of three-way
// bf16 splits (W's parts precomputed, the operand split exactly in registers), which
// keeps FP32 accuracy while running on the matrix cores concurrently with the VALU
// (FP32 MFMA on gfx950 shares the vector pipeline and gains nothing over it). A wave owns
// 32 paths = 2 N-tiles of 16 (the portfolio kernel's shape is a launch choice). The
// accumulator layout of a 16x16 tile -- lane l holds rows
// 4 (l / 16) + r, r = 0..3, of column l % 16 -- is used directly as the next block's B
// operand (W's columns are ordered to match), so the curve never leaves registers. The
// shocks of the block are generated on the VALU, one step per lane group per quad, and
// each 32-entry K chunk enters the matrix cores as soon as it is complete. One
// workgroup of 8 waves shares W's fragments in LDS.
//   ref_paths_mfma_kernel: reference semantics (up to 7 M-tiles): the block integrals dA are
//     Kahan-added into (A, C); numeraire slot b+1 is recorded after block b.
//   pf_paths_mfma_kernel: portfolio semantics (4 M-tiles): the curve is checkpointed at
//     every date and the short rate's left-point sum (row 51) extends the log bank account.
//   pf_value_mfma_kernel: the portfolio's netting-set values at one date, C = a - gamma P
//     over 16-set x 16-path tiles (sets as M, paths as N, periods as K).
//   pf_adjoint_seed_mfma_kernel: the portfolio adjoint's seeds, V = a - gamma P and
//     dP = -gamma^T u, the second product fed from the first's accumulators.
//   pf_value_stats_mfma_kernel: the same values at every date of a workgroup's paths,
//     reduced to EE sums, per-path CVA moments and the PFE's bracket counting where they
//     are produced (no exposures in HBM; opt-in).
#include "hjm_common.h"
#include "hjm_kernels.h"
#include "hjm_mfma_plan.h"
#include "hjm_rng.h"
#include "hjm_select.h"

#include <utility>

namespace hjm {

typedef float f4_t __attribute__((ext_vector_type(4)));
typedef __bf16 bf8_t __attribute__((ext_vector_type(8)));
typedef uint32_t u4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f4_t mfma_bf16(u4_t a, u4_t b, f4_t c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8_t, a),
                                                   __builtin_bit_cast(bf8_t, b), c, 0, 0, 0);
}

// Exact three-way split of 8 FP32 values into packed bf16 parts (h + m + l = x): h keeps
// the top 8 significant bits, m the next 8, l the last 8 (each difference is exact).
struct Split8 {
    u4_t h, m, l;
};

__device__ __forceinline__ Split8 split8(const float* x)
{
#pragma clang fp contract(off)
    uint32_t hb[8], mb[8], lb[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const uint32_t u = __builtin_bit_cast(uint32_t, x[i]);
        hb[i] = u & 0xffff0000u;
        const float r1 = x[i] - __builtin_bit_cast(float, hb[i]);
        mb[i] = __builtin_bit_cast(uint32_t, r1) & 0xffff0000u;
        const float r2 = r1 - __builtin_bit_cast(float, mb[i]);
        lb[i] = __builtin_bit_cast(uint32_t, r2);
    }
    Split8 s;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        // dword j = element 2j (low half) | element 2j+1 (high half)
        s.h[j] = __builtin_amdgcn_perm(hb[2 * j + 1], hb[2 * j], 0x07060302u);
        s.m[j] = __builtin_amdgcn_perm(mb[2 * j + 1], mb[2 * j], 0x07060302u);
        s.l[j] = __builtin_amdgcn_perm(lb[2 * j + 1], lb[2 * j], 0x07060302u);
    }
    return s;
}

// reference kernel: 8 waves of 2 N-tiles (16 paths each); the portfolio kernel's shape
// is a template choice (MfmaShape)
constexpr int kNT = 2;
constexpr int kMfmaBlock = 512;
constexpr int kMfmaPathsPerBlock = kMfmaBlock / kWave * kNT * 16;

// last quad whose values complete shock chunk s (values 8 s .. 8 s + 7, three per quad)
__host__ __device__ constexpr int shock_chunk_last_quad(int s)
{
    return (8 * s + 7) / 3 < kMfmaQuads - 1 ? (8 * s + 7) / 3 : kMfmaQuads - 1;
}

// Issue plan: the matrix-core work of the block is spread over the quads that generate
// shocks, so that their RNG (VALU) runs beside a share of the MFMAs: chunk c's M-tiles
// [m0, m1) go to quad q. Quads 3, 7 and 11 make no Philox call (their shocks come from
// calls made earlier) and get no MFMA work; curve chunks 0, 1 are ready at the start,
// shock chunk s once its last shock's call is made (quads 1, 4, 6, 9, 12); shock chunk 4
// completes with the last quad and runs after it.
struct IssueSlot {
    int chunk, m0, m1;
};
template <int MT>
__host__ __device__ constexpr IssueSlot issue_slot(int q, int half)
{
    const int split = (MT + 1) / 2;
    (void)half;
    switch (q) {
    case 0: return {0, 0, split};
    case 1: return {0, split, MT};
    case 2: return {1, 0, MT};
    case 4: return {2, 0, split};
    case 5: return {2, split, MT};
    case 6: return {3, 0, split};
    case 8: return {3, split, MT};
    case 9: return {4, 0, split};
    case 10: return {4, split, MT};
    case 12: return {5, 0, MT};
    default: return {-1, 0, 0};
    }
}

// f(integral_constant<int, Q>) for Q in the sequence, in order (compile-time indices).
template <class F, int... Q>
__device__ __forceinline__ void for_each_index(F& f, std::integer_sequence<int, Q...>)
{
    (f(std::integral_constant<int, Q>{}), ...);
}

// One 50-step block: acc = W x [fin; shocks of steps b*50 .. b*50+49; 1]. kStore: also
// write the block's normals to zs[step][factor][path] (FP32) for a later reverse sweep.
// kTwoAcc: the five small split products (Wh xm, Wm xh, Wh xl, Wl xh, Wm xm) accumulate
// apart from Wh xh and join it once at the end. The matrix cores drop the low bits of
// products far below the accumulator (measured: C = 1 plus 32 products of 2^-27 returns 1;
// small positive contributions are lost, small negative ones kept), which biases a
// single accumulator holding the curve (FP32 vs FP64: EE 1.4e-6 -> 3.2e-7 on the 10y
// precision workload with two). It costs 4 registers per tile and N-tile.
template <int MT, int NT, bool kStore, bool kInterleave, bool kTwoAcc>
__device__ __forceinline__ void mfma_block(const u4_t* __restrict__ lds_w, const float (&fin)[NT][16],
                                           f4_t (&acc)[MT][NT], const MfmaGrid& g,
                                           const int64_t (&path)[NT], const bool (&live)[NT],
                                           int b, int grp, int lane, float* __restrict__ zs)
{
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int n = 0; n < NT; ++n) acc[m][n] = f4_t{0.f, 0.f, 0.f, 0.f};
    // W's fragments are the same in every block: an opaque offset keeps their LDS reads
    // inside the block instead of hoisted (hundreds of registers) out of the block loop
    uint32_t off = (uint32_t)lane;
    asm volatile("" : "+v"(off));
    constexpr int MT2 = kTwoAcc ? MT : 1;
    f4_t acc2[MT2][NT];
#pragma unroll
    for (int m = 0; m < MT2; ++m)
#pragma unroll
        for (int n = 0; n < NT; ++n) acc2[m][n] = f4_t{0.f, 0.f, 0.f, 0.f};
    // M-tiles [m0, m1) of chunk c, less the structurally zero curve chunks (the reference
    // and portfolio maps share tiles 0..3, whose zeros agree; c and m are constants once
    // unrolled)
    auto tiles = [&](int c, int m0, int m1, const Split8 (&xs)[NT]) {
#pragma unroll
        for (int m = m0; m < m1; ++m) {
            if (mfma_chunk_zero(m, c, false)) continue;
            const u4_t* a = lds_w + ((m * kMfmaChunks + c) * kMfmaParts) * kWave + off;
            const u4_t ah = a[0], am = a[kWave], al = a[2 * kWave];
#pragma unroll
            for (int n = 0; n < NT; ++n) {
                f4_t v;
                if constexpr (kTwoAcc) {
                    acc[m][n] = mfma_bf16(ah, xs[n].h, acc[m][n]);
                    v = acc2[m][n];
                } else {
                    v = mfma_bf16(ah, xs[n].h, acc[m][n]);
                }
                v = mfma_bf16(ah, xs[n].m, v);
                v = mfma_bf16(am, xs[n].h, v);
                v = mfma_bf16(ah, xs[n].l, v);
                v = mfma_bf16(al, xs[n].h, v);
                v = mfma_bf16(am, xs[n].m, v);
                if constexpr (kTwoAcc) acc2[m][n] = v;
                else acc[m][n] = v;
            }
        }
    };
    // split operands of the 7 chunks: curve chunks from fin (registers 8 c .. 8 c + 7),
    // shock chunk s from values 8 s .. 8 s + 7 of z
    Split8 xs[kMfmaChunks][NT];
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int n = 0; n < NT; ++n) xs[c][n] = split8(&fin[n][8 * c]);
    // shocks: value v = 3 q + k of lane group grp is factor k of step 4 q + grp; step 50
    // (lane group 2, q = 12) carries the constant 1 of the drift in factor 0.
    float z[NT][8 * (kMfmaChunks - 2)];
#pragma unroll
    for (int n = 0; n < NT; ++n) z[n][3 * kMfmaQuads] = 0.f;
    auto quad = [&](auto qc) {
        constexpr int q = decltype(qc)::value;
        // steps 4 q + grp < 50 for every lane group up to q = 11; only the last quad pads
        constexpr bool last = 4 * q + 3 >= kMfmaStride;
        const int t = 4 * q + grp;
        const bool in = !last || t < kMfmaStride;
        // the lane group's class stream: shocks 3 q .. 3 q + 2 of this quad come from calls
        // (3 q) / 4 .. (3 q + 2) / 4; the calls this quad needs first are made here (quads
        // 3, 7 and 11 make none: 10 calls per block)
        constexpr int chi = (3 * q + 2) / 4;
        constexpr int clo = q == 0 ? 0 : (3 * q - 1) / 4 + 1;
#pragma unroll
        for (int n = 0; n < NT; ++n) {
#pragma unroll
            for (int c = clo; c <= chi; ++c) {
                float v[4];
                stream_call(g.seed_lo, g.seed_hi, (uint32_t)(g.path0 + path[n]), (uint32_t)b,
                            (uint32_t)grp, (uint32_t)c, v);
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    if (4 * c + e < 3 * kMfmaQuads) z[n][4 * c + e] = v[e];
            }
            if constexpr (last) {
                z[n][3 * q + 0] = in ? z[n][3 * q + 0] : (t == kMfmaStride ? 1.f : 0.f);
                z[n][3 * q + 1] = in ? z[n][3 * q + 1] : 0.f;
                z[n][3 * q + 2] = in ? z[n][3 * q + 2] : 0.f;
            }
            if constexpr (kStore) {
                if (in && live[n]) {
                    float* o = zs + (int64_t)(b * kMfmaStride + t) * 3 * g.paths + path[n];
                    o[0] = z[n][3 * q + 0];
                    o[g.paths] = z[n][3 * q + 1];
                    o[2 * g.paths] = z[n][3 * q + 2];
                }
            }
        }
#pragma unroll
        for (int s = 0; s < kMfmaChunks - 2; ++s)
            if (q == shock_chunk_last_quad(s))
#pragma unroll
                for (int n = 0; n < NT; ++n) xs[2 + s][n] = split8(&z[n][8 * s]);
        // this quad's share of the matrix work (its operands are complete); it runs on
        // the matrix cores beside the RNG of this and the other waves
        constexpr IssueSlot is = issue_slot<MT>(q, 0);
        if constexpr (is.chunk >= 0) {
            tiles(is.chunk, is.m0, is.m1, xs[is.chunk]);
            if constexpr (kInterleave) {
                // within the quad: one MFMA, then a share of the RNG's VALU instructions
#pragma unroll
                for (int i = 0; i < (is.m1 - is.m0) * NT * 6; ++i) {
                    __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
                    __builtin_amdgcn_sched_group_barrier(0x002, 7, 0);
                }
            }
        }
        // each quad is its own scheduling region (keeps the group solver's problems small)
        if constexpr (kInterleave) __builtin_amdgcn_sched_barrier(0);
    };
    for_each_index(quad, std::make_integer_sequence<int, kMfmaQuads>{});
    tiles(kMfmaChunks - 1, 0, MT, xs[kMfmaChunks - 1]);
    if constexpr (kTwoAcc) {
#pragma unroll
        for (int m = 0; m < MT; ++m)
#pragma unroll
            for (int n = 0; n < NT; ++n) acc[m][n] += acc2[m][n];
    }
}

// Reference semantics. MT = the live M-tiles of the block map: the curve (tiles 0..3,
// rows 0..50) and the integral rows 51 + j up to the last recorded slot j = ndates (a
// tile past it holds tenors whose numeraire is never recorded: ref_mfma_tiles).
template <bool kStore, int MT>
__global__ __launch_bounds__(kMfmaBlock) void ref_paths_mfma_kernel(const MfmaGrid g,
                                                                const u4_t* __restrict__ wfrag,
                                                                double* __restrict__ num_x,
                                                                double* __restrict__ num_y,
                                                                float* __restrict__ zs)
{
    constexpr int kInt = MT - 3;  // M-tiles holding integral rows (tile 3 holds both kinds)
    extern __shared__ u4_t lds_w[];
    for (int t = threadIdx.x; t < MT * kMfmaChunks * kMfmaParts * kWave; t += kMfmaBlock)
        lds_w[t] = wfrag[t];
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
    // Kahan (A, C) of the rows of tiles 3 .. MT-1 (dA_j at row 51 + j)
    float Ak[kNT][4 * kInt], Ck[kNT][4 * kInt];
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
        for (int m = 3; m < MT; ++m)
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
        f4_t acc[MT][kNT];
        // two accumulators where registers allow (up to 5 live tiles: 239 VGPRs)
        mfma_block<MT, kNT, kStore, true, (MT <= 5)>(lds_w, fin, acc, g, path, live, b, grp, lane, zs);
        // integrals: Kahan-add the block's dA (tiles 3 .. MT-1); curve rows become next input
#pragma unroll
        for (int n = 0; n < kNT; ++n) {
#pragma unroll
            for (int m = 3; m < MT; ++m)
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
        // slot b+1: x = curve row b+1, integral = row 51 + b + 1 (uniform positions; the
        // launcher's MT covers row 51 + ndates)
        const int slot = b + 1;
        const int Rx = slot, Ra = kTenors + slot;
        const int gx = (Rx % 16) / 4, ix = 4 * (Rx / 16) + Rx % 4;
        const int ga = (Ra % 16) / 4, ia = 4 * (Ra / 16 - 3) + Ra % 4;
#pragma unroll
        for (int n = 0; n < kNT; ++n) {
            float xv = 0.f, av = 0.f, cv = 0.f;
#pragma unroll
            for (int t2 = 0; t2 < 16; ++t2)
                if (t2 == ix) xv = fin[n][t2];
#pragma unroll
            for (int t2 = 0; t2 < 4 * kInt; ++t2)
                if (t2 == ia) {
                    av = Ak[n][t2];
                    cv = Ck[n][t2];
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
// A workgroup of WAVES waves of NT N-tiles simulates path group blockIdx.x of WAVES x NT x
// 16 paths; kPersist: then every gridDim.x-th group after it, with a grid smaller than the
// number of groups (the block map is staged in LDS once per workgroup). The group loop is
// a separate instantiation: around the date loop it costs the compiler's allocation
// about 50 VGPRs.
template <bool kStore, int NT, int WAVES, bool kPersist>
__global__ __launch_bounds__(WAVES * kWave) void pf_paths_mfma_kernel(const MfmaGrid g,
                                                                  const u4_t* __restrict__ wfrag,
                                                                  float* __restrict__ curves,
                                                                  double* __restrict__ logb,
                                                                  float* __restrict__ zs)
{
    constexpr int kThreads = WAVES * kWave, kGroup = WAVES * NT * 16;
    extern __shared__ u4_t lds_w[];
    for (int t = threadIdx.x; t < kMfmaPfFragElems / 8; t += kThreads) lds_w[t] = wfrag[t];
    __syncthreads();

    const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    const int grp = lane >> 4, col = lane & 15;
    const int64_t P = g.paths;
    const int64_t groups = (P + kGroup - 1) / kGroup;
    // row 51 (the short-rate sum) is register 4 * 3 + 3 of lane group 0
    constexpr int kSumReg = 4 * (kTenors / 16) + kTenors % 4;
    const bool sum_lane = grp == (kTenors % 16) / 4;
    for (int64_t gi = blockIdx.x; gi < groups; gi += kPersist ? gridDim.x : groups) {
        const int64_t base = gi * kGroup + (int64_t)wv * NT * 16;
        int64_t path[NT];
        bool live[NT];
        float fin[NT][16];
        double lnB[NT];
        // opaque per group: the initial curve is selected again in every group rather than
        // held in 16 registers across the loop
        int lg = grp;
        asm volatile("" : "+v"(lg));
#pragma unroll
        for (int n = 0; n < NT; ++n) {
            path[n] = base + 16 * n + col;
            live[n] = path[n] < P;
            lnB[n] = 0.0;
#pragma unroll
            for (int m = 0; m < 4; ++m)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int R = 16 * m + 4 * lg + r;
                    fin[n][4 * m + r] = (R < kTenors) ? g.spot[R < kTenors ? R : 0] : 0.f;
                }
        }
        for (int e = 0;; ++e) {
#pragma unroll
            for (int n = 0; n < NT; ++n) {
                if (!live[n]) continue;
#pragma unroll
                for (int m = 0; m < 4; ++m)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int R = 16 * m + 4 * grp + r;
                        if (R < kTenors)
                            curves[((int64_t)e * kTenors + R) * P + path[n]] = fin[n][4 * m + r];
                    }
                if (sum_lane) logb[(int64_t)e * P + path[n]] = lnB[n];
            }
            if (e + 1 >= g.ndates) break;
            f4_t acc[kMfmaPfMTiles][NT];
            mfma_block<kMfmaPfMTiles, NT, kStore, false, true>(lds_w, fin, acc, g, path, live, e, grp, lane,
                                                               zs);
#pragma unroll
            for (int n = 0; n < NT; ++n) {
#pragma unroll
                for (int m = 0; m < 4; ++m)
#pragma unroll
                    for (int r = 0; r < 4; ++r) fin[n][4 * m + r] = acc[m][n][r];
                lnB[n] += (double)fin[n][kSumReg] * g.dt;
            }
        }
    }
}

// The portfolio path kernel's instantiated launch shapes {N-tiles per wave, waves}: {2, 8},
// {2, 12}, {1, 12}, {1, 16} (the block map's 84 KB of LDS allow one workgroup per CU;
// 2 N-tiles take about 164 VGPRs, at most 3 waves per SIMD).
constexpr int kPathShapes = 4;
template <bool kStore, bool kPersist>
struct PfPathsKernels {
    template <int NT, int W>
    static const void* fn() { return (const void*)pf_paths_mfma_kernel<kStore, NT, W, kPersist>; }
};

hipError_t mfma_kernels_prepare()
{
    const void* fns[8 + 4 * kPathShapes] = {
        (const void*)ref_paths_mfma_kernel<false, 4>, (const void*)ref_paths_mfma_kernel<false, 5>,
        (const void*)ref_paths_mfma_kernel<false, 6>, (const void*)ref_paths_mfma_kernel<false, 7>,
        (const void*)ref_paths_mfma_kernel<true, 4>,  (const void*)ref_paths_mfma_kernel<true, 5>,
        (const void*)ref_paths_mfma_kernel<true, 6>,  (const void*)ref_paths_mfma_kernel<true, 7>};
    int lds[8 + 4 * kPathShapes];
    for (int i = 0; i < 8; ++i) lds[i] = ref_mfma_lds_bytes(4 + i % 4);
    int k = 8;
    auto add = [&](const void* f) {
        fns[k] = f;
        lds[k++] = 2 * kMfmaPfFragElems;
    };
    auto add_shapes = [&](auto kernels) {
        using K = decltype(kernels);
        add(K::template fn<2, 8>());
        add(K::template fn<2, 12>());
        add(K::template fn<1, 12>());
        add(K::template fn<1, 16>());
    };
    add_shapes(PfPathsKernels<false, false>{});
    add_shapes(PfPathsKernels<true, false>{});
    add_shapes(PfPathsKernels<false, true>{});
    add_shapes(PfPathsKernels<true, true>{});
    for (int i = 0; i < k; ++i) {
        const hipError_t e = hipFuncSetAttribute(fns[i], hipFuncAttributeMaxDynamicSharedMemorySize, lds[i]);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_ref_paths_mfma(const MfmaGrid& g, const uint16_t* wfrag, double* nx, double* ny,
                                 float* zs, hipStream_t st)
{
    if (g.ndates < 1 || g.ndates > kTenors - 1) return hipErrorInvalidValue;
    const dim3 grid((unsigned)((g.paths + kMfmaPathsPerBlock - 1) / kMfmaPathsPerBlock));
    const u4_t* w = reinterpret_cast<const u4_t*>(wfrag);
    const int mt = ref_mfma_tiles(g.ndates);
    const size_t lds = (size_t)ref_mfma_lds_bytes(mt);
#define HJM_REF_LAUNCH(MT)                                                                      \
    (zs ? ref_paths_mfma_kernel<true, MT><<<grid, kMfmaBlock, lds, st>>>(g, w, nx, ny, zs)      \
        : ref_paths_mfma_kernel<false, MT><<<grid, kMfmaBlock, lds, st>>>(g, w, nx, ny, zs))
    switch (mt) {
    case 4: HJM_REF_LAUNCH(4); break;
    case 5: HJM_REF_LAUNCH(5); break;
    case 6: HJM_REF_LAUNCH(6); break;
    default: HJM_REF_LAUNCH(7); break;
    }
#undef HJM_REF_LAUNCH
    return hipGetLastError();
}

bool pf_path_shape_ok(int tiles, int waves)
{
    return (tiles == 2 && (waves == 8 || waves == 12)) || (tiles == 1 && (waves == 12 || waves == 16));
}

template <bool kStore, bool kPersist>
static hipError_t launch_pf_paths_shape(const MfmaGrid& g, const u4_t* w, float* curves, double* logb,
                                        float* zs, int tiles, int waves, int64_t blocks, hipStream_t st)
{
    const dim3 grid((unsigned)blocks), block((unsigned)(waves * kWave));
    const size_t lds = 2 * kMfmaPfFragElems;
    if (tiles == 2 && waves == 8)
        pf_paths_mfma_kernel<kStore, 2, 8, kPersist><<<grid, block, lds, st>>>(g, w, curves, logb, zs);
    else if (tiles == 2 && waves == 12)
        pf_paths_mfma_kernel<kStore, 2, 12, kPersist><<<grid, block, lds, st>>>(g, w, curves, logb, zs);
    else if (tiles == 1 && waves == 12)
        pf_paths_mfma_kernel<kStore, 1, 12, kPersist><<<grid, block, lds, st>>>(g, w, curves, logb, zs);
    else if (tiles == 1 && waves == 16)
        pf_paths_mfma_kernel<kStore, 1, 16, kPersist><<<grid, block, lds, st>>>(g, w, curves, logb, zs);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_pf_paths_mfma(const MfmaGrid& g, const uint16_t* wfrag, float* curves,
                                double* logb, float* zs, const MfmaShape& shape, hipStream_t st)
{
    const int tiles = shape.tiles ? shape.tiles : 2, waves = shape.waves ? shape.waves : 8;
    if (!pf_path_shape_ok(tiles, waves) || shape.max_blocks < 0) return hipErrorInvalidValue;
    const u4_t* w = reinterpret_cast<const u4_t*>(wfrag);
    const int64_t group = (int64_t)waves * tiles * 16;
    const int64_t groups = (g.paths + group - 1) / group;
    if (shape.max_blocks > 0 && shape.max_blocks < groups) {
        const int64_t b = shape.max_blocks;
        return zs ? launch_pf_paths_shape<true, true>(g, w, curves, logb, zs, tiles, waves, b, st)
                  : launch_pf_paths_shape<false, true>(g, w, curves, logb, zs, tiles, waves, b, st);
    }
    return zs ? launch_pf_paths_shape<true, false>(g, w, curves, logb, zs, tiles, waves, groups, st)
              : launch_pf_paths_shape<false, false>(g, w, curves, logb, zs, tiles, waves, groups, st);
}

// Netting-set exposures of one date on the matrix cores (FP32 plans):
//   E[set][path] = max(a_set - sum_k gamma[set][k] P_k(path), 0) D(path)
// as C = A B + a, A = -gamma (16 sets x 32 periods per fragment, its three bf16 parts
// precomputed per date), B = the paths' bond prices P_k, split exactly in registers, six
// split products per tile as in the path kernels. Lane l of a wave owns path 16 n + l % 16
// of N-tile n and periods 8 (l / 16) .. + 7 of each of the two 32-period chunks: it builds
// the trapezoid sums of its two 8-period segments from the curve checkpoint and takes the
// sums of the segments before them from the other lane groups (shuffles), then P_k =
// exp(-I_k). The accumulator layout (lane l: sets 4 (l / 16) + r of path l % 16) is stored
// directly: per store, 16 consecutive paths of one set. An M-tile whose sets have no live
// period past 32 skips the second chunk. The valuation and the fused valuation-statistics
// kernel below share these helpers, so both produce the same exposures bit for bit.

// The bond prices of one N-tile's path (column l % 16) at date e, split for the B operand
// of the two 32-period chunks, and its discount factor (from lane group `src_grp`'s load,
// shuffled to the others by the caller).
__device__ __forceinline__ void value_bonds(const PfValueParams& v, int e, int64_t p,
                                            const float* __restrict__ curves, int grp, int col,
                                            Split8 (&bs)[2], float* __restrict__ pk = nullptr)
{
    int64_t P = v.paths;
    // opaque per call: a caller's date loop neither holds the 18 rows' offsets (8 grp + i) P
    // in VGPRs nor their addresses
    asm volatile("" : "+s"(P));
    const float hw = (float)(0.5 * v.dtau), dtf = (float)v.dtau;
    const float* fc = curves + (int64_t)e * kTenors * P + p;
    asm volatile("" : "+v"(fc));
    // segment 0: periods k = 8 grp + 1 .. 8 grp + 8 (tenors 8 grp .. 8 grp + 8 <= 32);
    // segment 1: k = 33 + 8 grp .. 40 + 8 grp (tenors capped at the last, 50)
    float fa[9], fb[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) fa[i] = fc[(int64_t)(8 * grp + i) * P];
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const int j = 32 + 8 * grp + i;
        fb[i] = fc[(int64_t)(j < kTenors - 1 ? j : kTenors - 1) * P];
    }
    float ca[8], cb[8];
    float s = 0.f;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        s += hw * (fa[i] + fa[i + 1]);
        ca[i] = s;
    }
    s = 0.f;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const int k = 33 + 8 * grp + i;
        s += k <= kTenors - 1 ? hw * (fb[i] + fb[i + 1]) : dtf * fb[i];
        cb[i] = s;
    }
    float t0[4], t1[4];
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        t0[g] = __shfl(ca[7], col + 16 * g, kWave);
        t1[g] = __shfl(cb[7], col + 16 * g, kWave);
    }
    float offa = 0.f, offb = ((t0[0] + t0[1]) + t0[2]) + t0[3];
#pragma unroll
    for (int g = 0; g < 3; ++g) {
        offa += g < grp ? t0[g] : 0.f;
        offb += g < grp ? t1[g] : 0.f;
    }
    float pa[8], pb[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        pa[i] = expf(-(offa + ca[i]));
        pb[i] = 32 + 8 * grp + i < kMaxPeriods ? expf(-(offb + cb[i])) : 0.f;
    }
    bs[0] = split8(pa);
    bs[1] = split8(pb);
    // pk (or null): the unsplit prices of the lane's periods 8 grp + i and 32 + 8 grp + i
    if (pk) {
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            pk[8 * grp + i] = pa[i];
            pk[32 + 8 * grp + i] = pb[i];
        }
    }
}

// One M-tile's values for NT N-tiles: acc[n] = a + (-gamma) P over the chunks the tile uses.
// (A second accumulator for the small split products, as in the path kernels, changed the
// per-set CVA error by 2 % and cost the valuation a wave per SIMD: not used here.)
template <int NT>
__device__ __forceinline__ void value_tile(const u4_t* __restrict__ fr0, int lane, int tk, f4_t a4,
                                           const Split8 (&bs)[NT][2], f4_t (&acc)[NT])
{
#pragma unroll
    for (int n = 0; n < NT; ++n) acc[n] = a4;
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        if (tk > 32 * c) {
            const u4_t* fr = fr0 + (c * kMfmaParts) * kWave + lane;
            const u4_t ah = fr[0], am = fr[kWave], al = fr[2 * kWave];
#pragma unroll
            for (int n = 0; n < NT; ++n) {
                f4_t x = acc[n];
                x = mfma_bf16(ah, bs[n][c].h, x);
                x = mfma_bf16(ah, bs[n][c].m, x);
                x = mfma_bf16(am, bs[n][c].h, x);
                x = mfma_bf16(ah, bs[n][c].l, x);
                x = mfma_bf16(al, bs[n][c].h, x);
                x = mfma_bf16(am, bs[n][c].m, x);
                acc[n] = x;
            }
        }
    }
}

__device__ __forceinline__ float exposure_of(float V, float disc) { return V > 0.f ? V * disc : 0.f; }

// Exposures of paths [0, count) at the date blockIdx.y (rows [date][set], row stride
// v.paths). A workgroup of WAVES waves of NT N-tiles values WAVES x NT x 16 paths; the
// date's -gamma fragments pass through LDS `stage` M-tiles at a time (dynamic LDS). Rows of
// sets whose swaps have all matured (v.kmax 0) are not written: the plan keeps them 0.
// filter (or null): value only the M-tiles holding a set whose PFE row at this date the
// counting pass left unresolved (bracket_row_resolved false for rank fk and fcap
// candidate slots); a workgroup with none returns at once.
constexpr int kValMaxStage = 8;  // M-tiles staged at a time, at most

template <int NT, int WAVES>
__global__ __launch_bounds__(WAVES * kWave) void pf_value_mfma_kernel(const PfValueParams v, int mtiles,
                                                                    int stage, int64_t count,
                                                                    const float* __restrict__ curves,
                                                                    const double* __restrict__ logb,
                                                                    const u4_t* __restrict__ gfrag,
                                                                    const float* __restrict__ apad,
                                                                    const int* __restrict__ tilek,
                                                                    const SelRowState* __restrict__ filter,
                                                                    int64_t fk, int64_t fcap,
                                                                    float* __restrict__ exposures)
{
    constexpr int kThreads = WAVES * kWave, kPaths = WAVES * NT * 16;
    extern __shared__ u4_t sfrag[];
    __shared__ uint32_t s_need[kMaxPfMTiles];
    __shared__ uint8_t s_live[HJM_MAX_SETS];
    __shared__ int s_any;
    const int e = blockIdx.y;
    const int64_t P = v.paths;
    const int G = v.nsets;
    const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
    const int col = lane % 16, grp = lane / 16;
    const int64_t pbase = (int64_t)blockIdx.x * kPaths + wave * (NT * 16);
    if (filter) {
        if (threadIdx.x == 0) s_any = 0;
        for (int m = threadIdx.x; m < mtiles; m += kThreads) s_need[m] = 0;
        __syncthreads();
        for (int g = threadIdx.x; g < G; g += kThreads)
            if (!bracket_row_resolved(filter[(int64_t)e * G + g], fk, fcap)) {
                s_need[g / 16] = 1;
                s_any = 1;
            }
        __syncthreads();
        if (!s_any) return;
    }
    // discount factors of the wave's N-tiles: lane group j loads N-tile j's (j < NT) and
    // shares it with the other lane groups
    float disc[NT];
    {
        const int64_t p = pbase + 16 * (grp < NT ? grp : 0) + col;
        const float d = (float)exp(-logb[(int64_t)e * P + (p < count ? p : 0)]);
#pragma unroll
        for (int n = 0; n < NT; ++n) disc[n] = __shfl(d, col + 16 * n, kWave);
    }
    Split8 bs[NT][2];
#pragma unroll
    for (int n = 0; n < NT; ++n) {
        const int64_t p = pbase + 16 * n + col;
        value_bonds(v, e, p < count ? p : 0, curves, grp, col, bs[n]);
    }
    for (int g = threadIdx.x; g < G; g += kThreads) s_live[g] = v.kmax[(int64_t)e * G + g] > 0;
    for (int m0 = 0; m0 < mtiles; m0 += stage) {
        const int mt = mtiles - m0 < stage ? mtiles - m0 : stage;
        __syncthreads();
        const u4_t* src = gfrag + ((int64_t)e * mtiles + m0) * kValueFragsPerTile * kWave;
        for (int t = threadIdx.x; t < mt * kValueFragsPerTile * kWave; t += kThreads) sfrag[t] = src[t];
        __syncthreads();
        for (int mm = 0; mm < mt; ++mm) {
            const int m = m0 + mm;
            if (filter && !s_need[m]) continue;
            const int tk = tilek[(int64_t)e * mtiles + m];
            if (tk == 0) continue;   // every set of the tile has matured
            const f4_t a4 = *reinterpret_cast<const f4_t*>(apad + ((int64_t)e * mtiles + m) * 16 + 4 * grp);
            f4_t acc[NT];
            value_tile<NT>(sfrag + (mm * kValueFragsPerTile) * kWave, lane, tk, a4, bs, acc);
#pragma unroll
            for (int n = 0; n < NT; ++n) {
                const int64_t p = pbase + 16 * n + col;
                if (p < count) {
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int g = 16 * m + 4 * grp + r;
                        if (g < G && s_live[g])
                            exposures[((int64_t)e * G + g) * P + p] = exposure_of(acc[n][r], disc[n]);
                    }
                }
            }
        }
    }
}

bool pf_value_shape_ok(int tiles, int waves)
{
    return (tiles == 2 || tiles == 4) && (waves == 4 || waves == 8);
}

hipError_t launch_pf_value_mfma(const PfValueParams& v, const float* curves, const double* logb,
                                const uint16_t* gfrag, const float* apad, const int* tilek,
                                float* ex, const MfmaShape& shape, hipStream_t st, int64_t count,
                                const SelRowState* filter, int64_t fk, int64_t fcap)
{
    if (v.paths <= 0 || v.nsets <= 0 || v.ndates <= 0) return hipErrorInvalidValue;
    const int mtiles = value_mtiles(v.nsets);
    if (mtiles > kMaxPfMTiles) return hipErrorInvalidValue;
    const int tiles = shape.tiles ? shape.tiles : 4, waves = shape.waves ? shape.waves : 4;
    if (!pf_value_shape_ok(tiles, waves) || shape.stage < 0 || shape.stage > kValMaxStage)
        return hipErrorInvalidValue;
    const int stage = std::min(shape.stage ? shape.stage : kValMaxStage, mtiles);
    if (count < 0 || count > v.paths) count = v.paths;
    if (count == 0) return hipSuccess;
    const int64_t paths_per_block = (int64_t)waves * tiles * 16;
    const dim3 grid((unsigned)((count + paths_per_block - 1) / paths_per_block), (unsigned)v.ndates);
    const dim3 block((unsigned)(waves * kWave));
    const size_t lds = (size_t)stage * kValueFragsPerTile * kWave * sizeof(u4_t);
    const u4_t* gf = reinterpret_cast<const u4_t*>(gfrag);
#define HJM_VALUE_LAUNCH(NT, W)                                                                  \
    pf_value_mfma_kernel<NT, W><<<grid, block, lds, st>>>(v, mtiles, stage, count, curves, logb, gf, \
                                                          apad, tilek, filter, fk, fcap, ex)
    if (tiles == 4 && waves == 4) HJM_VALUE_LAUNCH(4, 4);
    else if (tiles == 4 && waves == 8) HJM_VALUE_LAUNCH(4, 8);
    else if (tiles == 2 && waves == 4) HJM_VALUE_LAUNCH(2, 4);
    else HJM_VALUE_LAUNCH(2, 8);
#undef HJM_VALUE_LAUNCH
    return hipGetLastError();
}

// the sum over the wave's four rows of 16 lanes, in every lane
template <typename T>
__device__ __forceinline__ T cross_row_sum(T x)
{
    x += __shfl_xor(x, 16, kWave);
    return x + __shfl_xor(x, 32, kWave);
}

// ---- seeds of the portfolio adjoint (FP32 plans) ---------------------------------------
// Per path and date, from the curve checkpoint (as pf_adjoint_seed_kernel):
//   V = a - Gamma P (the exposure pass's product, sets as rows), u_g = w_g D 1[V_g > 0],
//   beta = -sum_g w_g max(V_g, 0) D,  dP = -Gamma^T u,  iota_k = -P_k dP_k,
//   gbar_l = dtau/2 (S_{l+1} [l <= 49] + S_l [l >= 1]) + [l = 50] dtau sum_{k>50} (k-50) iota_k,
//   S_m = sum_{k >= m} iota_k.
// Both products run on the matrix cores (bf16x6 split products). The second takes u
// straight from the first's accumulators: its K chunk c holds sets 16 (2c + i/4) + 4 g +
// i % 4 in lane group g's slot i, which are the registers i % 4 of set tiles 2c and 2c + 1
// (build_adjoint_fragments orders -Gamma^T the same way), and its output rows are the
// periods 16 pm + 4 g + r. A workgroup of 8 waves values 128 paths at one date
// (blockIdx.y), then the next of its kSeedGroups groups; -Gamma and -Gamma^T of 128 sets
// at a time are staged in LDS (once per workgroup up to 128 sets), the bond prices go
// through LDS to the lanes that hold the matching dP, and each lane then forms the suffix
// sums of its path and stores the tenors 13 g .. 13 g + 12 of gbar.
constexpr int kSeedWaves = 8;
constexpr int kSeedBlock = kSeedWaves * kWave;
constexpr int kSeedPaths = kSeedWaves * 16;     // one N-tile per wave
constexpr int kSeedTiles = 8;                   // set tiles staged at a time
constexpr int kPeriodTiles = 4;                 // 64 period rows of the second product
constexpr int kSeedGroups = 8;                  // path groups per workgroup (staging amortised)

__global__ __launch_bounds__(kSeedBlock) void pf_adjoint_seed_mfma_kernel(
    const PfValueParams v, int mtiles, const float* __restrict__ curves,
    const double* __restrict__ logb, const u4_t* __restrict__ gfrag, const u4_t* __restrict__ tfrag,
    const float* __restrict__ apad, const int* __restrict__ tilek, const double* __restrict__ w,
    float* __restrict__ gbar, float* __restrict__ beta, double* __restrict__ beta_partials)
{
    __shared__ u4_t s1[kSeedTiles * kValueFragsPerTile * kWave];
    __shared__ u4_t s2[kPeriodTiles * (kSeedTiles / 2) * kMfmaParts * kWave];
    // per path: bond prices, then iota; rows of 65 floats so the 16 paths of a wave fall
    // in distinct LDS banks
    __shared__ float scr[kSeedWaves][16][65];
    __shared__ float s_w[16 * kSeedTiles];
    __shared__ int s_tk[kSeedTiles];
    __shared__ double s_b[kSeedWaves];
    const int e = blockIdx.y;
    const int64_t P = v.paths;
    const int G = v.nsets, nd = v.ndates;
    const int nch = (mtiles + 1) / 2;
    const bool one_group = mtiles <= kSeedTiles;   // fragments staged once per workgroup
    const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
    const int col = lane % 16, grp = lane / 16;
    const float hw = (float)(0.5 * v.dtau), dtf = (float)v.dtau;
    auto stage = [&](int m0, int nt) {
        const int c0 = m0 / 2, ncg = (nt + 1) / 2;
        const u4_t* src = gfrag + ((int64_t)e * mtiles + m0) * kValueFragsPerTile * kWave;
        for (int t = threadIdx.x; t < nt * kValueFragsPerTile * kWave; t += kSeedBlock) s1[t] = src[t];
        for (int t = threadIdx.x; t < kPeriodTiles * ncg * kMfmaParts * kWave; t += kSeedBlock) {
            const int q = t / (ncg * kMfmaParts * kWave), rest = t % (ncg * kMfmaParts * kWave);
            s2[(q * (kSeedTiles / 2)) * kMfmaParts * kWave + rest] =
                tfrag[(((int64_t)e * kPeriodTiles + q) * nch + c0) * kMfmaParts * kWave + rest];
        }
        for (int gl = threadIdx.x; gl < 16 * kSeedTiles; gl += kSeedBlock) {
            const int g = 16 * m0 + gl;
            s_w[gl] = g < G ? (float)w[(int64_t)g * nd + e] : 0.f;
        }
        if (threadIdx.x < kSeedTiles)
            s_tk[threadIdx.x] = (int)threadIdx.x < nt ? tilek[(int64_t)e * mtiles + m0 + threadIdx.x] : 0;
    };
    if (one_group) {
        stage(0, mtiles);
        __syncthreads();
    }
    double bsum = 0.0;   // the wave's sum of beta over its paths of all groups
    for (int pg = 0; pg < kSeedGroups; ++pg) {
        const int64_t p = ((int64_t)blockIdx.x * kSeedGroups + pg) * kSeedPaths + wave * 16 + col;
        const bool live = p < P;
        const int64_t pc = live ? p : 0;
        const float D = (float)exp(-logb[(int64_t)e * P + pc]);
        Split8 bs[1][2];
        value_bonds(v, e, pc, curves, grp, col, bs[0], &scr[wave][col][0]);
        f4_t dp[kPeriodTiles];
#pragma unroll
        for (int q = 0; q < kPeriodTiles; ++q) dp[q] = f4_t{0.f, 0.f, 0.f, 0.f};
        float b = 0.f;
        for (int m0 = 0; m0 < mtiles; m0 += kSeedTiles) {
            const int nt = mtiles - m0 < kSeedTiles ? mtiles - m0 : kSeedTiles;
            const int ncg = (nt + 1) / 2;
            if (!one_group) {
                __syncthreads();   // the previous group's fragments are no longer read
                stage(m0, nt);
                __syncthreads();
            }
            // V of the group's set tiles, u and beta where V is produced
            float u[kSeedTiles][4];
#pragma unroll
            for (int m = 0; m < kSeedTiles; ++m) {
                if (m < nt) {
                    const f4_t a4 =
                        *reinterpret_cast<const f4_t*>(apad + ((int64_t)e * mtiles + m0 + m) * 16 + 4 * grp);
                    f4_t acc[1];
                    value_tile<1>(s1 + (m * kValueFragsPerTile) * kWave, lane, s_tk[m], a4, bs, acc);
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const float V = acc[0][r], wg = s_w[16 * m + 4 * grp + r];
                        const bool pos = V > 0.f;
                        u[m][r] = pos ? wg * D : 0.f;
                        b -= pos ? wg * V * D : 0.f;
                    }
                } else {
#pragma unroll
                    for (int r = 0; r < 4; ++r) u[m][r] = 0.f;
                }
            }
            // dP += (-Gamma^T) u over the group's K chunks
#pragma unroll
            for (int cc = 0; cc < kSeedTiles / 2; ++cc) {
                if (cc < ncg) {
                    float x[8];
#pragma unroll
                    for (int i = 0; i < 8; ++i) x[i] = u[2 * cc + i / 4][i % 4];
                    const Split8 bu = split8(x);
#pragma unroll
                    for (int q = 0; q < kPeriodTiles; ++q) {
                        const u4_t* fr = s2 + ((q * (kSeedTiles / 2) + cc) * kMfmaParts) * kWave + lane;
                        const u4_t ah = fr[0], am = fr[kWave], al = fr[2 * kWave];
                        f4_t y = dp[q];
                        y = mfma_bf16(ah, bu.h, y);
                        y = mfma_bf16(ah, bu.m, y);
                        y = mfma_bf16(am, bu.h, y);
                        y = mfma_bf16(ah, bu.l, y);
                        y = mfma_bf16(al, bu.h, y);
                        y = mfma_bf16(am, bu.m, y);
                        dp[q] = y;
                    }
                }
            }
        }
        // beta of the lane's path: its sets' terms summed over the four lane groups
        b = cross_row_sum(b);
        if (grp == 0 && live) beta[(int64_t)e * P + p] = b;
        bsum += wave_sum((grp == 0 && live) ? (double)b : 0.0);
        // iota at the lane's dP rows (periods 16 q + 4 grp + r), over the prices in LDS (the
        // wave's own rows: its LDS accesses stay in program order)
#pragma unroll
        for (int q = 0; q < kPeriodTiles; ++q) {
            float* row = &scr[wave][col][16 * q + 4 * grp];
#pragma unroll
            for (int r = 0; r < 4; ++r) row[r] = -row[r] * dp[q][r];
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        // suffix sums S_k (k = 1 .. 60 is period index k - 1) and this lane's tenors of gbar
        const float* io = &scr[wave][col][0];
        float S[kTenors + 1];   // S[l] = S_l for l = 1 .. 51
        float run = 0.f, ext = 0.f;
#pragma unroll
        for (int k = kMaxPeriods; k >= 1; --k) {
            const float iota = io[k - 1];
            run += iota;
            if (k <= kTenors) S[k] = run;
            if (k > kTenors - 1) ext += (float)(k - (kTenors - 1)) * iota;
        }
        if (live) {
            float* gb = gbar + (int64_t)e * kTenors * P + p;
#pragma unroll
            for (int l = 0; l < kTenors; ++l) {
                if (l / 13 != grp) continue;
                float val = 0.f;
                if (l <= kTenors - 2) val += hw * S[l + 1];
                if (l >= 1) val += hw * S[l];
                if (l == kTenors - 1) val += dtf * ext;
                gb[(int64_t)l * P] = val;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();   // the next group's prices overwrite these rows
    }
    if (lane == 0) s_b[wave] = bsum;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
#pragma unroll
        for (int ww = 0; ww < kSeedWaves; ++ww) t += s_b[ww];
        beta_partials[(int64_t)blockIdx.x * nd + e] = t;
    }
}

int64_t pf_adjoint_seed_mfma_blocks(int64_t paths)
{
    return (paths + kSeedPaths * kSeedGroups - 1) / (kSeedPaths * kSeedGroups);
}

hipError_t launch_pf_adjoint_seed_mfma(const PfValueParams& v, const float* curves, const double* logb,
                                       const uint16_t* gfrag, const uint16_t* tfrag, const float* apad,
                                       const int* tilek, const double* w, float* gbar, float* beta,
                                       double* beta_partials, hipStream_t st)
{
    const int mtiles = value_mtiles(v.nsets);
    if (v.paths <= 0 || v.nsets <= 0 || v.ndates <= 0 || mtiles > kMaxPfMTiles) return hipErrorInvalidValue;
    const dim3 grid((unsigned)pf_adjoint_seed_mfma_blocks(v.paths), (unsigned)v.ndates);
    pf_adjoint_seed_mfma_kernel<<<grid, kSeedBlock, 0, st>>>(
        v, mtiles, curves, logb, reinterpret_cast<const u4_t*>(gfrag), reinterpret_cast<const u4_t*>(tfrag),
        apad, tilek, w, gbar, beta, beta_partials);
    return hipGetLastError();
}

// ---- fused valuation and statistics (FP32 plans) -------------------------------------
// The exposure pass's values, computed as it computes them (sets as the matrix rows, a
// wave's paths as the columns: lane l holds sets 16 n + 4 (l / 16) + r of path l % 16 of
// set tile n), reduced where they are produced so that the exposures never go to HBM. A
// workgroup of kFuseWaves waves x kFuseNT N-tiles (128 paths) takes a group of up to 128
// sets (blockIdx.y) through every date in turn: per date the bond prices once, then the
// group's set tiles, whose -gamma fragments pass through LDS kFuseStage tiles at a time.
// Per set tile and date:
//   * EE numerators: the 4 sets' sums over a wave's paths (in-lane over its N-tiles, then
//     over the 16 lanes of the row by a transposing reduction), then the waves in order in
//     FP64 at the date's flush: partials[set][block][date];
//   * the per-path CVA sample sum_e w_ge E_pge in registers (FP32) across the dates; its
//     sum and square sum per set at the end: partials[set][block][ndates + 0, 1];
//   * with sel: the counting pass of the PFE's sampled-bracket selection for every row
//     (date, set) -- entries below the bracket (or the zeros), counted four sets to a
//     packed word; entries inside it counted and staged in LDS (kFuseSlots per set and
//     date; more flags the row overflowed, so its selection values it again, whole), then
//     one claim per row and workgroup and the copy to the row's candidate buffer.
// The per-date arrays are double-buffered, so the flush of date e - 1 (one thread per set)
// runs beside the staging of date e. Sums are in fixed order and counts are integers:
// results do not depend on scheduling (the staged candidates' order does, which the
// selection does not see).
constexpr int kFuseWaves = 4;
constexpr int kFuseBlock = kFuseWaves * kWave;
constexpr int kFuseNT = 2;                                // N-tiles (16 paths) per wave
constexpr int kFusePaths = kFuseWaves * kFuseNT * 16;     // 128
constexpr int kFuseTiles = 8;                             // set tiles per group
constexpr int kFuseSets = 16 * kFuseTiles;                // 128 sets per group
constexpr int kFuseStage = 4;                             // set tiles' fragments in LDS at a time
constexpr int kFuseSlots = 16;                            // staged inside entries per set and date

struct FuseDate {
    float w[kFuseSets];                       // CVA weight of the set at the date (0 past the last set)
    uint32_t lo[kFuseSets], hi[kFuseSets];    // PFE bracket (FP32 bits)
    uint32_t nlo[kFuseSets], nin[kFuseSets];  // entries below / inside it
    float ee[kFuseWaves][kFuseSets];          // per-wave EE sums
    float cand[kFuseSets][kFuseSlots];        // staged inside entries
    int tk[kFuseTiles];                       // the set tile's largest live period
};

// Sums of v[0..3] over the 16 lanes of a row; lane c of the row ends with the sum of
// v[c / 4] (two exchanges of halves, then two plain steps: 5 shuffles instead of 16).
template <typename T>
__device__ __forceinline__ T row_sum4(const T (&v)[4], int col)
{
    const bool h8 = col & 8, h4 = col & 4;
    T s0 = h8 ? v[2] : v[0], s1 = h8 ? v[3] : v[1];
    s0 += __shfl_xor(h8 ? v[0] : v[2], 8, 16);
    s1 += __shfl_xor(h8 ? v[1] : v[3], 8, 16);
    T u = h4 ? s1 : s0;
    u += __shfl_xor(h4 ? s0 : s1, 4, 16);
    u += __shfl_xor(u, 2, 16);
    return u + __shfl_xor(u, 1, 16);
}

// 2 waves per SIMD: the per-path CVA holds 64 registers beside the valuation's (at 3 the
// kernel spills 372 bytes per lane; with one N-tile per wave it runs slower)
__global__ __launch_bounds__(kFuseBlock) __attribute__((amdgpu_waves_per_eu(2))) void pf_value_stats_mfma_kernel(
    const PfValueParams v, int mtiles, const float* __restrict__ curves,
    const double* __restrict__ logb, const u4_t* __restrict__ gfrag, const float* __restrict__ apad,
    const int* __restrict__ tilek, const double* __restrict__ w, double* __restrict__ partials,
    SelRowState* __restrict__ sel, float* __restrict__ cand, int64_t cap, int slots)
{
    __shared__ u4_t sfrag[kFuseStage * kValueFragsPerTile * kWave];
    __shared__ FuseDate sd[2];
    const int64_t P = v.paths;
    const int G = v.nsets, nd = v.ndates, ncol = nd + 2;
    const int m0 = blockIdx.y * kFuseTiles, g0 = 16 * m0;   // the group's first set tile, set
    const int nt = mtiles - m0 < kFuseTiles ? mtiles - m0 : kFuseTiles;
    const int ng = G - g0 < kFuseSets ? G - g0 : kFuseSets;
    const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
    const int col = lane % 16, grp = lane / 16;
    const int blocks = (int)gridDim.x;
    const int64_t pbase = (int64_t)blockIdx.x * kFusePaths + wave * (kFuseNT * 16);
    bool live[kFuseNT];
#pragma unroll
    for (int n = 0; n < kFuseNT; ++n) live[n] = pbase + 16 * n + col < P;
    float cvap[kFuseTiles][kFuseNT][4];
#pragma unroll
    for (int t = 0; t < kFuseTiles; ++t)
#pragma unroll
        for (int n = 0; n < kFuseNT; ++n)
#pragma unroll
            for (int r = 0; r < 4; ++r) cvap[t][n][r] = 0.f;

    auto stage_date = [&](int e, FuseDate& d) {
        for (int gl = threadIdx.x; gl < kFuseSets; gl += kFuseBlock) {
            const bool in = gl < ng;
            const int64_t rid = (int64_t)e * G + g0 + gl;
            d.w[gl] = in ? (float)w[(int64_t)(g0 + gl) * nd + e] : 0.f;
            d.lo[gl] = (in && sel) ? (uint32_t)sel[rid].lo : 0u;
            d.hi[gl] = (in && sel) ? (uint32_t)sel[rid].hi : 0u;
            d.nlo[gl] = 0u;
            d.nin[gl] = 0u;
        }
        if (threadIdx.x < kFuseTiles)
            d.tk[threadIdx.x] = (int)threadIdx.x < nt ? tilek[(int64_t)e * mtiles + m0 + threadIdx.x] : 0;
    };
    // one thread per set: EE partial, counts, claim, candidates of date e
    auto flush = [&](int e, const FuseDate& d) {
        for (int gl = threadIdx.x; gl < ng; gl += kFuseBlock) {
            double t = 0.0;
#pragma unroll
            for (int ww = 0; ww < kFuseWaves; ++ww) t += (double)d.ee[ww][gl];
            partials[((int64_t)(g0 + gl) * blocks + blockIdx.x) * ncol + e] = t;
            if (!sel) continue;
            const int64_t rid = (int64_t)e * G + g0 + gl;
            const uint32_t lo = d.lo[gl], hi = d.hi[gl], nlo = d.nlo[gl], nin = d.nin[gl];
            const uint32_t lo1 = lo ? lo : 1u;
            const bool collect = hi >= lo1 && lo != hi;
            if (nlo) atomicAdd(lo ? &sel[rid].below : &sel[rid].zin, nlo);
            const bool over = collect && nin > (uint32_t)slots;
            const uint32_t take = (collect && !over) ? nin : 0u;
            const uint32_t base = nin ? sel_count_and_claim(&sel[rid], nin, take) : 0u;
            if (over) {
                atomicOr(&sel[rid].claim, kSelOverflow);
            } else if (take) {
                for (uint32_t i = 0; i < nin; ++i) {
                    const int64_t slot = (int64_t)base + i;
                    if (slot < cap) cand[rid * cap + slot] = d.cand[gl][i];
                }
            }
        }
    };

    for (int e = 0; e < nd; ++e) {
        FuseDate& d = sd[e & 1];
        stage_date(e, d);
        if (e > 0) flush(e - 1, sd[(e - 1) & 1]);
        // discount factors (lane group n loads N-tile n's and shares it) and bond prices
        float disc[kFuseNT];
        {
            const int64_t p = pbase + 16 * (grp < kFuseNT ? grp : 0) + col;
            const float df = (float)exp(-logb[(int64_t)e * P + (p < P ? p : 0)]);
#pragma unroll
            for (int n = 0; n < kFuseNT; ++n) disc[n] = __shfl(df, col + 16 * n, kWave);
        }
        Split8 bs[kFuseNT][2];
#pragma unroll
        for (int n = 0; n < kFuseNT; ++n) {
            const int64_t p = pbase + 16 * n + col;
            value_bonds(v, e, p < P ? p : 0, curves, grp, col, bs[n]);
            // one N-tile's curve rows and partial sums live at a time
            __builtin_amdgcn_sched_barrier(0);
        }
        // the group's set tiles in pieces of kFuseStage; each piece a compile-time index, so
        // the per-path CVA registers are addressed by constants
        auto piece = [&](auto pc) {
            constexpr int s0 = decltype(pc)::value * kFuseStage;
            if (s0 >= nt) return;
            const int st = nt - s0 < kFuseStage ? nt - s0 : kFuseStage;
            // the previous piece's fragments are consumed (and, at s0 = 0, date e's arrays
            // are staged and date e - 1's were flushed from the other buffer)
            __syncthreads();
            const u4_t* src = gfrag + ((int64_t)e * mtiles + m0 + s0) * kValueFragsPerTile * kWave;
            for (int t = threadIdx.x; t < st * kValueFragsPerTile * kWave; t += kFuseBlock) sfrag[t] = src[t];
            __syncthreads();
#pragma unroll
            for (int mm = 0; mm < kFuseStage; ++mm) {
                const int n = s0 + mm;   // set tile within the group
                if (mm >= st) break;
                const f4_t a4 = *reinterpret_cast<const f4_t*>(
                    apad + ((int64_t)e * mtiles + m0 + n) * 16 + 4 * grp);
                f4_t acc[kFuseNT];
                // a matured tile (tk 0) makes no products; its zeros are still counted
                value_tile<kFuseNT>(sfrag + (mm * kValueFragsPerTile) * kWave, lane, d.tk[n], a4, bs, acc);
                float es[4];
                uint32_t below = 0;   // four 8-bit counts (at most 32 each)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int gl = 16 * n + 4 * grp + r;
                    const float wg = d.w[gl];
                    const uint32_t lo = d.lo[gl], hi = d.hi[gl];
                    const uint32_t lo1 = lo ? lo : 1u;   // inside: non-zero and >= lo
                    const bool any_in = sel && hi >= lo1;
                    const bool collect = any_in && lo != hi;
                    es[r] = 0.f;
#pragma unroll
                    for (int k = 0; k < kFuseNT; ++k) {
                        const float E = live[k] ? exposure_of(acc[k][r], disc[k]) : 0.f;
                        cvap[n][k][r] = __builtin_fmaf(wg, E, cvap[n][k][r]);
                        es[r] += E;
                        const uint32_t b = __builtin_bit_cast(uint32_t, E);
                        below += (live[k] && b < lo1) ? (1u << (8 * r)) : 0u;
                        if (any_in && live[k] && b - lo1 <= hi - lo1) {
                            const uint32_t pos = atomicAdd(&d.nin[gl], 1u);
                            if (collect && pos < (uint32_t)slots) d.cand[gl][pos] = E;
                        }
                    }
                }
                const float et = row_sum4(es, col);
                const int gr = 16 * n + 4 * grp + col / 4;   // the set lane col's sums are of
                if ((col & 3) == 0) d.ee[wave][gr] = et;
                if (sel) {
                    uint32_t cb = below;
                    cb += __shfl_xor(cb, 8, 16);
                    cb += __shfl_xor(cb, 4, 16);
                    cb += __shfl_xor(cb, 2, 16);
                    cb += __shfl_xor(cb, 1, 16);
                    const uint32_t nlo = (cb >> (8 * (col / 4))) & 0xffu;
                    if ((col & 3) == 0 && nlo) atomicAdd(&d.nlo[gr], nlo);
                }
                // each set tile is its own scheduling region: interleaving the tiles would
                // keep their operands and accumulators live together
                __builtin_amdgcn_sched_barrier(0);
            }
        };
        for_each_index(piece, std::make_integer_sequence<int, kFuseTiles / kFuseStage>{});
        // date e's arrays are complete before the next iteration flushes them
        __syncthreads();
    }
    flush(nd - 1, sd[(nd - 1) & 1]);
    // per-path CVA moments per set: in-lane over the N-tiles, then the row, then the
    // waves in order
    __syncthreads();
    double* s_cva = reinterpret_cast<double*>(sfrag);   // [wave][set][2]
#pragma unroll
    for (int n = 0; n < kFuseTiles; ++n) {
        if (n >= nt) break;
        double c1[4], c2[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            c1[r] = 0.0;
            c2[r] = 0.0;
#pragma unroll
            for (int k = 0; k < kFuseNT; ++k) {
                const double c = (double)cvap[n][k][r];
                c1[r] += c;
                c2[r] += c * c;
            }
        }
        const double t1 = row_sum4(c1, col), t2 = row_sum4(c2, col);
        const int gr = 16 * n + 4 * grp + col / 4;
        if ((col & 3) == 0) {
            s_cva[(wave * kFuseSets + gr) * 2] = t1;
            s_cva[(wave * kFuseSets + gr) * 2 + 1] = t2;
        }
        __builtin_amdgcn_sched_barrier(0);   // one tile's FP64 values live at a time
    }
    __syncthreads();
    for (int t = threadIdx.x; t < 2 * ng; t += kFuseBlock) {
        const int gl = t / 2, q = t % 2;
        double s = 0.0;
#pragma unroll
        for (int ww = 0; ww < kFuseWaves; ++ww) s += s_cva[(ww * kFuseSets + gl) * 2 + q];
        partials[((int64_t)(g0 + gl) * blocks + blockIdx.x) * ncol + nd + q] = s;
    }
}

int64_t pf_fused_blocks(int64_t paths) { return (paths + kFusePaths - 1) / kFusePaths; }

hipError_t launch_pf_value_stats_mfma(const PfValueParams& v, const float* curves, const double* logb,
                                      const uint16_t* gfrag, const float* apad, const int* tilek,
                                      const double* w, double* partials, SelRowState* sel,
                                      float* cand, int64_t cap, int stage_slots, hipStream_t st)
{
    const int slots = stage_slots > 0 && stage_slots < kFuseSlots ? stage_slots : kFuseSlots;
    const int mtiles = value_mtiles(v.nsets);
    if (v.paths <= 0 || v.nsets <= 0 || mtiles > kMaxPfMTiles) return hipErrorInvalidValue;
    const dim3 grid((unsigned)pf_fused_blocks(v.paths), (unsigned)((mtiles + kFuseTiles - 1) / kFuseTiles));
    pf_value_stats_mfma_kernel<<<grid, kFuseBlock, 0, st>>>(v, mtiles, curves, logb,
                                                            reinterpret_cast<const u4_t*>(gfrag),
                                                            apad, tilek, w, partials, sel, cand, cap,
                                                            slots);
    return hipGetLastError();
}

}  // namespace hjm
