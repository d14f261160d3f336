// hjm_mfma_kernels.hip -- FP32 path kernels on the matrix cores.
//
// Each 50-step block is the affine map [f_out; integral rows] = W [f_in; phi; 1]
// (hjm_mfma_plan.h), computed with v_mfma_f32_16x16x32_bf16 as six products of three-way
// bf16 splits (W's parts precomputed, the operand split exactly in registers), which
// keeps FP32 accuracy while running on the matrix cores concurrently with the VALU
// (FP32 MFMA on gfx950 shares the vector pipeline and gains nothing over it). A wave owns
// 32 paths = 2 N-tiles of 16. The accumulator layout of a 16x16 tile -- lane l holds rows
// 4 (l / 16) + r, r = 0..3, of column l % 16 -- is used directly as the next block's B
// operand (W's columns are ordered to match), so the curve never leaves registers. The
// shocks of the block are generated on the VALU, one step per lane group per quad, and
// each 32-entry K chunk enters the matrix cores as soon as it is complete. One
// workgroup of 8 waves shares W's fragments in LDS.
//   ref_paths_mfma_kernel: reference semantics (7 M-tiles): the block integrals dA are
//     Kahan-added into (A, C); numeraire slot b+1 is recorded after block b.
//   pf_paths_mfma_kernel: portfolio semantics (4 M-tiles): the curve is checkpointed at
//     every date and the short rate's left-point sum (row 51) extends the log bank account.
//   pf_value_mfma_kernel: the portfolio's netting-set values at one date, C = a - gamma P
//     over 16-set x 16-path tiles (sets as M, paths as N, periods as K).
#include "hjm_common.h"
#include "hjm_kernels.h"
#include "hjm_mfma_plan.h"
#include "hjm_rng.h"

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

constexpr int kNT = 2;                       // N-tiles (16 paths each) per wave
constexpr int kMfmaBlock = 512;              // 8 waves
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
template <int MT, bool kStore>
__device__ __forceinline__ void mfma_block(const u4_t* __restrict__ lds_w, const float (&fin)[kNT][16],
                                           f4_t (&acc)[MT][kNT], const MfmaGrid& g,
                                           const int64_t (&path)[kNT], const bool (&live)[kNT],
                                           int b, int grp, int lane, float* __restrict__ zs)
{
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int n = 0; n < kNT; ++n) acc[m][n] = f4_t{0.f, 0.f, 0.f, 0.f};
    // W's fragments are the same in every block: an opaque offset keeps their LDS reads
    // inside the block instead of hoisted (hundreds of registers) out of the block loop
    uint32_t off = (uint32_t)lane;
    asm volatile("" : "+v"(off));
    auto tiles = [&](int c, int m0, int m1, const Split8 (&xs)[kNT]) {
#pragma unroll
        for (int m = m0; m < m1; ++m) {
            const u4_t* a = lds_w + ((m * kMfmaChunks + c) * kMfmaParts) * kWave + off;
            const u4_t ah = a[0], am = a[kWave], al = a[2 * kWave];
#pragma unroll
            for (int n = 0; n < kNT; ++n) {
                f4_t v = acc[m][n];
                v = mfma_bf16(ah, xs[n].h, v);
                v = mfma_bf16(ah, xs[n].m, v);
                v = mfma_bf16(am, xs[n].h, v);
                v = mfma_bf16(ah, xs[n].l, v);
                v = mfma_bf16(al, xs[n].h, v);
                v = mfma_bf16(am, xs[n].m, v);
                acc[m][n] = v;
            }
        }
    };
    // split operands of the 7 chunks: curve chunks from fin (registers 8 c .. 8 c + 7),
    // shock chunk s from values 8 s .. 8 s + 7 of z
    Split8 xs[kMfmaChunks][kNT];
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int n = 0; n < kNT; ++n) xs[c][n] = split8(&fin[n][8 * c]);
    // shocks: value v = 3 q + k of lane group grp is factor k of step 4 q + grp; step 50
    // (lane group 2, q = 12) carries the constant 1 of the drift in factor 0.
    float z[kNT][8 * (kMfmaChunks - 2)];
#pragma unroll
    for (int n = 0; n < kNT; ++n) z[n][3 * kMfmaQuads] = 0.f;
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
        for (int n = 0; n < kNT; ++n) {
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
                for (int n = 0; n < kNT; ++n) xs[2 + s][n] = split8(&z[n][8 * s]);
        // this quad's share of the matrix work (its operands are complete); it runs on
        // the matrix cores beside the RNG of this and the other waves
        constexpr IssueSlot is = issue_slot<MT>(q, 0);
        if constexpr (is.chunk >= 0) {
            tiles(is.chunk, is.m0, is.m1, xs[is.chunk]);
            if constexpr (MT == kMfmaMTiles) {
                // within the quad: one MFMA, then a share of the RNG's VALU instructions
#pragma unroll
                for (int i = 0; i < (is.m1 - is.m0) * kNT * 6; ++i) {
                    __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
                    __builtin_amdgcn_sched_group_barrier(0x002, 7, 0);
                }
            }
        }
        // each quad is its own scheduling region (keeps the group solver's problems small)
        if constexpr (MT == kMfmaMTiles) __builtin_amdgcn_sched_barrier(0);
    };
    for_each_index(quad, std::make_integer_sequence<int, kMfmaQuads>{});
    tiles(kMfmaChunks - 1, 0, MT, xs[kMfmaChunks - 1]);
}

template <bool kStore>
__global__ __launch_bounds__(kMfmaBlock) void ref_paths_mfma_kernel(const MfmaGrid g,
                                                                const u4_t* __restrict__ wfrag,
                                                                double* __restrict__ num_x,
                                                                double* __restrict__ num_y,
                                                                float* __restrict__ zs)
{
    extern __shared__ u4_t lds_w[];
    for (int t = threadIdx.x; t < kMfmaFragElems / 8; t += kMfmaBlock) lds_w[t] = wfrag[t];
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
        mfma_block<kMfmaMTiles, kStore>(lds_w, fin, acc, g, path, live, b, grp, lane, zs);
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
template <bool kStore>
__global__ __launch_bounds__(kMfmaBlock) void pf_paths_mfma_kernel(const MfmaGrid g,
                                                               const u4_t* __restrict__ wfrag,
                                                               float* __restrict__ curves,
                                                               double* __restrict__ logb,
                                                               float* __restrict__ zs)
{
    extern __shared__ u4_t lds_w[];
    for (int t = threadIdx.x; t < kMfmaPfFragElems / 8; t += kMfmaBlock) lds_w[t] = wfrag[t];
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
        mfma_block<kMfmaPfMTiles, kStore>(lds_w, fin, acc, g, path, live, e, grp, lane, zs);
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
    const void* fns[4] = {(const void*)ref_paths_mfma_kernel<false>, (const void*)ref_paths_mfma_kernel<true>,
                          (const void*)pf_paths_mfma_kernel<false>, (const void*)pf_paths_mfma_kernel<true>};
    const int lds[4] = {2 * kMfmaFragElems, 2 * kMfmaFragElems, 2 * kMfmaPfFragElems,
                        2 * kMfmaPfFragElems};
    for (int i = 0; i < 4; ++i) {
        const hipError_t e = hipFuncSetAttribute(fns[i], hipFuncAttributeMaxDynamicSharedMemorySize, lds[i]);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_ref_paths_mfma(const MfmaGrid& g, const uint16_t* wfrag, double* nx, double* ny,
                                 float* zs, hipStream_t st)
{
    const dim3 grid((unsigned)((g.paths + kMfmaPathsPerBlock - 1) / kMfmaPathsPerBlock));
    const u4_t* w = reinterpret_cast<const u4_t*>(wfrag);
    if (zs)
        ref_paths_mfma_kernel<true><<<grid, kMfmaBlock, 2 * kMfmaFragElems, st>>>(g, w, nx, ny, zs);
    else
        ref_paths_mfma_kernel<false><<<grid, kMfmaBlock, 2 * kMfmaFragElems, st>>>(g, w, nx, ny, zs);
    return hipGetLastError();
}

hipError_t launch_pf_paths_mfma(const MfmaGrid& g, const uint16_t* wfrag, float* curves,
                                double* logb, float* zs, hipStream_t st)
{
    const dim3 grid((unsigned)((g.paths + kMfmaPathsPerBlock - 1) / kMfmaPathsPerBlock));
    const u4_t* w = reinterpret_cast<const u4_t*>(wfrag);
    if (zs)
        pf_paths_mfma_kernel<true><<<grid, kMfmaBlock, 2 * kMfmaPfFragElems, st>>>(g, w, curves, logb, zs);
    else
        pf_paths_mfma_kernel<false><<<grid, kMfmaBlock, 2 * kMfmaPfFragElems, st>>>(g, w, curves, logb, zs);
    return hipGetLastError();
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
// directly: per store, 16 consecutive paths of one set. A workgroup of 4 waves x 64 paths
// stages the date's A fragments in LDS, 8 M-tiles (128 sets) at a time; an M-tile whose
// sets have no live period past 32 skips the second chunk.
constexpr int kValBlock = 256;
constexpr int kValNT = 4;                                  // N-tiles (16 paths) per wave
constexpr int kValPaths = kValBlock / kWave * kValNT * 16; // paths per workgroup
constexpr int kValLdsTiles = 8;                            // M-tiles staged at a time

__global__ __launch_bounds__(kValBlock) void pf_value_mfma_kernel(const PfValueParams v, int mtiles,
                                                                const float* __restrict__ curves,
                                                                const double* __restrict__ logb,
                                                                const u4_t* __restrict__ gfrag,
                                                                const float* __restrict__ apad,
                                                                const int* __restrict__ tilek,
                                                                float* __restrict__ exposures)
{
    __shared__ u4_t sfrag[kValLdsTiles * kValueFragsPerTile * kWave];
    const int e = blockIdx.y;
    const int64_t P = v.paths;
    const int G = v.nsets;
    const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
    const int col = lane % 16, grp = lane / 16;
    const int64_t pbase = (int64_t)blockIdx.x * kValPaths + wave * (kValNT * 16);
    const float hw = (float)(0.5 * v.dtau), dtf = (float)v.dtau;
    // discount factor of N-tile grp's path, shared with the other lane groups below
    float disc[kValNT];
    {
        const int64_t p = pbase + 16 * grp + col;
        const float d = (float)exp(-logb[(int64_t)e * P + (p < P ? p : 0)]);
#pragma unroll
        for (int n = 0; n < kValNT; ++n) disc[n] = __shfl(d, col + 16 * n, kWave);
    }
    Split8 bs[kValNT][2];
#pragma unroll
    for (int n = 0; n < kValNT; ++n) {
        const int64_t p = pbase + 16 * n + col;
        const float* fc = curves + (int64_t)e * kTenors * P + (p < P ? p : 0);
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
        bs[n][0] = split8(pa);
        bs[n][1] = split8(pb);
    }
    for (int m0 = 0; m0 < mtiles; m0 += kValLdsTiles) {
        const int mt = mtiles - m0 < kValLdsTiles ? mtiles - m0 : kValLdsTiles;
        __syncthreads();
        const u4_t* src = gfrag + ((int64_t)e * mtiles + m0) * kValueFragsPerTile * kWave;
        for (int t = threadIdx.x; t < mt * kValueFragsPerTile * kWave; t += kValBlock) sfrag[t] = src[t];
        __syncthreads();
        for (int mm = 0; mm < mt; ++mm) {
            const int m = m0 + mm;
            const int tk = tilek[(int64_t)e * mtiles + m];
            const f4_t a4 = *reinterpret_cast<const f4_t*>(apad + ((int64_t)e * mtiles + m) * 16 + 4 * grp);
            f4_t acc[kValNT];
#pragma unroll
            for (int n = 0; n < kValNT; ++n) acc[n] = a4;
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                if (tk > 32 * c) {
                    const u4_t* fr = sfrag + (mm * kValueFragsPerTile + c * kMfmaParts) * kWave + lane;
                    const u4_t ah = fr[0], am = fr[kWave], al = fr[2 * kWave];
#pragma unroll
                    for (int n = 0; n < kValNT; ++n) {
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
#pragma unroll
            for (int n = 0; n < kValNT; ++n) {
                const int64_t p = pbase + 16 * n + col;
                if (p < P) {
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int g = 16 * m + 4 * grp + r;
                        const float V = acc[n][r];
                        if (g < G) exposures[((int64_t)e * G + g) * P + p] = V > 0.f ? V * disc[n] : 0.f;
                    }
                }
            }
        }
    }
}

hipError_t launch_pf_value_mfma(const PfValueParams& v, const float* curves, const double* logb,
                                const uint16_t* gfrag, const float* apad, const int* tilek,
                                float* ex, hipStream_t st)
{
    if (v.paths <= 0 || v.nsets <= 0 || v.ndates <= 0) return hipErrorInvalidValue;
    const int mtiles = value_mtiles(v.nsets);
    const dim3 grid((unsigned)((v.paths + kValPaths - 1) / kValPaths), (unsigned)v.ndates);
    pf_value_mfma_kernel<<<grid, kValBlock, 0, st>>>(v, mtiles, curves, logb,
                                                     reinterpret_cast<const u4_t*>(gfrag), apad,
                                                     tilek, ex);
    return hipGetLastError();
}

}  // namespace hjm
