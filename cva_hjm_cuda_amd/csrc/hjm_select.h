// hjm_select.h -- exact k-th order statistic of each row of non-negative values (PFE).
//
// MSD radix selection on the IEEE bit patterns (order-preserving for x >= 0), one
// workgroup of 1024 threads per row, digits of 11 bits (the last one narrower):
//   * each pass builds an LDS histogram of the next digit among the elements that match
//     the prefix fixed so far, then finds the digit holding rank k with a parallel scan
//     of the 2048 bins (two bins per thread: wave scans + a scan of the 16 wave totals);
//   * as soon as the chosen bin holds at most kSelCand elements they are copied to LDS
//     (one more pass over the row) and the remaining digits are resolved there, so a
//     64-bit row is streamed from HBM three or four times rather than six;
//   * exact zeros (exposures of out-of-the-money paths, often most of a row) are counted
//     per thread instead of by LDS atomics on one bin, and when a row holds at least k
//     zeros the answer is 0 after the first pass.
#pragma once
#include <math.h>
#include <stddef.h>
#include <stdint.h>

#include "hjm_common.h"

namespace hjm {

constexpr int kSelBlock = 1024;
constexpr int kSelBins = 2048;
constexpr int kSelCand = 4096;

template <typename T>
struct SelectBits;
template <>
struct SelectBits<double> {
    typedef uint64_t U;
    static constexpr int kBits = 64;
};
template <>
struct SelectBits<float> {
    typedef uint32_t U;
    static constexpr int kBits = 32;
};

// stage: a row of at most kSelCand entries goes to LDS first and every pass reads it
// there (the bracket selection's candidate rows).
template <typename T>
__device__ void select_kth_row(const T* __restrict__ row, int64_t n, int64_t k, double* out,
                               bool stage = false)
{
    typedef typename SelectBits<T>::U U;
    constexpr int NB = SelectBits<T>::kBits;
    constexpr int kWaves = kSelBlock / kWave;
    __shared__ uint32_t hist[kSelBins];
    __shared__ U cand[kSelCand];
    __shared__ uint32_t wtot[kWaves];
    __shared__ U s_prefix;
    __shared__ int64_t s_k;
    __shared__ uint32_t s_count, s_ncand, s_zeros;

    const int tid = threadIdx.x, lane = tid & (kWave - 1), wv = tid / kWave;
    if (tid == 0) {
        s_prefix = 0;
        s_k = k;
        s_zeros = 0;
    }
    bool collected = false;
    uint32_t ncand = 0;
    if (stage && n <= kSelCand) {
        // visible to every thread after the first pass's barrier
        for (int64_t t = tid; t < n; t += kSelBlock) cand[t] = __builtin_bit_cast(U, row[t]);
        collected = true;
        ncand = (uint32_t)n;
    }
    int known = 0;
    while (known < NB) {
        const int width = (NB - known) < 11 ? (NB - known) : 11;
        const int shift = NB - known - width;
        const U mask = (U)((1u << width) - 1u);
        for (int b = tid; b < kSelBins; b += kSelBlock) hist[b] = 0;
        __syncthreads();
        const U prefix = s_prefix;
        const int64_t kk = s_k;
        auto matches = [&](U bits) { return known == 0 || (bits >> (NB - known)) == (prefix >> (NB - known)); };
        if (!collected) {
            // zeros lie in bin 0 of every pass whose prefix they match: count them apart
            const bool zero_in = matches((U)0);
            uint32_t zc = 0;
            for (int64_t t = tid; t < n; t += kSelBlock) {
                const U bits = __builtin_bit_cast(U, row[t]);
                if (bits == 0) {
                    zc += zero_in ? 1u : 0u;
                } else if (matches(bits)) {
                    atomicAdd(&hist[(uint32_t)((bits >> shift) & mask)], 1u);
                }
            }
            for (int o = kWave / 2; o > 0; o >>= 1) zc += __shfl_xor(zc, o, kWave);
            if (lane == 0 && zc) atomicAdd(&hist[0], zc);
            if (known == 0 && lane == 0 && zc) atomicAdd(&s_zeros, zc);
        } else {
            for (uint32_t t = tid; t < ncand; t += kSelBlock) {
                const U bits = cand[t];
                if (matches(bits)) atomicAdd(&hist[(uint32_t)((bits >> shift) & mask)], 1u);
            }
        }
        __syncthreads();
        if (known == 0 && (int64_t)s_zeros >= kk) {
            // at least k exact zeros: the k-th smallest of a non-negative row is 0 (the
            // prefix is still 0 after the first pass's histogram)
            break;
        }
        // exclusive scan of the bins, two per thread
        const uint32_t h0 = hist[2 * tid], h1 = hist[2 * tid + 1];
        uint32_t incl = h0 + h1;
        for (int o = 1; o < kWave; o <<= 1) {
            const uint32_t v = __shfl_up(incl, o, kWave);
            if (lane >= o) incl += v;
        }
        if (lane == kWave - 1) wtot[wv] = incl;
        __syncthreads();
        uint32_t before = 0;
        for (int q = 0; q < wv; ++q) before += wtot[q];
        const int64_t excl = (int64_t)before + incl - (h0 + h1);
        if (excl < kk && kk <= excl + h0) {
            s_prefix = prefix | ((U)(2 * tid) << shift);
            s_k = kk - excl;
            s_count = h0;
        } else if (excl + h0 < kk && kk <= excl + h0 + h1) {
            s_prefix = prefix | ((U)(2 * tid + 1) << shift);
            s_k = kk - excl - h0;
            s_count = h1;
        }
        __syncthreads();
        known += width;
        if (known < NB && !collected && s_count <= (uint32_t)kSelCand) {
            // copy the elements of the chosen bin to LDS; later passes read only those
            if (tid == 0) s_ncand = 0;
            __syncthreads();
            const U pre = s_prefix;
            for (int64_t t = tid; t < n; t += kSelBlock) {
                const U bits = __builtin_bit_cast(U, row[t]);
                if ((bits >> (NB - known)) == (pre >> (NB - known))) {
                    const uint32_t pos = atomicAdd(&s_ncand, 1u);
                    cand[pos] = bits;
                }
            }
            __syncthreads();
            ncand = s_ncand;
            collected = true;
        }
    }
    if (tid == 0) *out = (double)__builtin_bit_cast(T, s_prefix);
}

}  // namespace hjm

namespace hjm {
template <typename T>
__global__ __launch_bounds__(kSelBlock) void select_kth_kernel(const T* __restrict__ x, int64_t n,
                                                               int64_t k, double* __restrict__ out)
{
    select_kth_row<T>(x + (int64_t)blockIdx.x * n, n, k, out + blockIdx.x);
}
}  // namespace hjm

// ---------------------------------------------------------------------------------------
// Sampled-bracket selection for long rows (PFE over many paths).
//
// The paths of a row are independent and identically distributed, so the first m entries
// of a row are a random sample of it. Three launches replace the three or four streaming
// passes of select_kth_row with one:
//   1. bracket_sample_kernel (one workgroup per row): bounds on the sample's order
//      statistics of ranks klo and khi around k m / n (6 binomial standard deviations
//      either side), by two radix passes in LDS; they become the bracket [lo, hi] (lo = 0
//      when klo = 1, hi = all ones when khi = m); the row's counters are cleared;
//   2. bracket_count_kernel (many workgroups per row): one pass over the row counting
//      below = #{x < lo}, the zeros inside the bracket (lo = 0) and inside = #{x != 0 in
//      [lo, hi]}, copying the latter to the row's candidate buffer (LDS-staged, one
//      global atomic per workgroup);
//   3. bracket_finish_kernel (one workgroup per row): when below < k <= below + zeros +
//      inside the answer is 0, lo (a bracket of one value) or the (k - below - zeros)-th
//      smallest candidate (select_kth_row over the candidates); otherwise -- the sample
//      missed, or the candidates overflowed their buffer -- select_kth_row over the whole
//      row. Every outcome is the exact k-th smallest element.

namespace hjm {

struct SelRowState {
    uint64_t lo, hi;            // bracket bit patterns
    uint32_t below, zin, nin;   // counts of the row (n < 2^32)
    uint32_t claim;             // candidate slots handed out so far
};

struct BracketPlan {
    bool use;       // false: select_kth_kernel over whole rows
    int32_t m;      // sample size (first m entries of a row)
    int32_t klo, khi;
    int64_t cap;    // candidate slots per row
    int32_t chunks; // counting workgroups per row
};

constexpr int kBracketSampleBytes = 64 * 1024;  // LDS of the sample
constexpr int kCountBlock = 256;
constexpr int kCountChunk = 8192;               // entries per counting workgroup
constexpr int kBracketMinRatio = 4;             // rows shorter than 4 m use select_kth_kernel

inline BracketPlan bracket_plan(int64_t n, int64_t k, int elem_bytes)
{
    BracketPlan bp{};
    const int64_t mmax = kBracketSampleBytes / elem_bytes;
    if (n < kBracketMinRatio * mmax || n >= (int64_t)1 << 32 || k < 1 || k > n) return bp;
    bp.use = true;
    bp.m = (int32_t)mmax;
    const double q = (double)k / (double)n;
    const double center = q * bp.m, sd = sqrt(bp.m * q * (1.0 - q)) + 1.0;
    const double lo = floor(center - 6.0 * sd) - 2.0, hi = ceil(center + 6.0 * sd) + 2.0;
    bp.klo = (int32_t)(lo < 1.0 ? 1.0 : lo);
    bp.khi = (int32_t)(hi > bp.m ? (double)bp.m : hi);
    const double expect = (double)(bp.khi - bp.klo + 1) / bp.m * (double)n;
    const int64_t cap = (int64_t)(2.0 * expect) + 8192;
    bp.cap = cap < n ? cap : n;
    bp.chunks = (int32_t)((n + kCountChunk - 1) / kCountChunk);
    return bp;
}

// bytes of the device workspace of a bracket selection over `rows` rows (0 when unused)
inline size_t bracket_workspace_bytes(int rows, const BracketPlan& bp, int elem_bytes)
{
    if (!bp.use || rows <= 0) return 0;
    const size_t head = ((size_t)rows * sizeof(SelRowState) + 255) & ~(size_t)255;
    return head + (size_t)rows * (size_t)bp.cap * (size_t)elem_bytes;
}

// The sample's order statistics need not be exact: any lo <= x_(k) <= hi is a valid
// bracket (the counting pass decides). Keys are taken relative to the smallest non-zero
// sample key, so two histogram passes of 11-bit digits spread over the sample's own range
// (no LDS-atomic pile-up on the few bins that hold a narrow distribution's exponent) and
// resolve the bracket to 2^-22 of that range; zeros are counted apart. lo is the lower
// edge of the bin holding rank klo, hi the upper edge of the one holding khi.
template <typename T>
__global__ __launch_bounds__(kSelBlock) void bracket_sample_kernel(const T* __restrict__ x, int64_t n,
                                                                   BracketPlan bp,
                                                                   SelRowState* __restrict__ state)
{
    typedef typename SelectBits<T>::U U;
    constexpr int NB = SelectBits<T>::kBits;
    constexpr int kMax = kBracketSampleBytes / (int)sizeof(U);
    constexpr int kWaves = kSelBlock / kWave;
    constexpr int kDigit = 11;
    __shared__ U keys[kMax];
    __shared__ uint32_t hist[2][kSelBins];
    __shared__ uint32_t wtot[2][kWaves];
    __shared__ U s_min[kWaves], s_max[kWaves];
    __shared__ uint32_t s_zero[kWaves];
    __shared__ U s_pre[2];
    __shared__ int64_t s_k[2];
    const int row = blockIdx.x, tid = threadIdx.x, lane = tid & (kWave - 1), wv = tid / kWave;
    const T* r = x + (int64_t)row * n;
    const int m = bp.m;
    U mn = (U)~(U)0, mx = 0;
    uint32_t zc = 0;
    for (int i = tid; i < m; i += kSelBlock) {
        const U key = __builtin_bit_cast(U, r[i]);
        keys[i] = key;
        if (key == 0) {
            ++zc;
        } else {
            mn = key < mn ? key : mn;
            mx = key > mx ? key : mx;
        }
    }
    for (int o = kWave / 2; o > 0; o >>= 1) {
        const U a = __shfl_xor(mn, o, kWave), c = __shfl_xor(mx, o, kWave);
        mn = a < mn ? a : mn;
        mx = c > mx ? c : mx;
        zc += __shfl_xor(zc, o, kWave);
    }
    if (lane == 0) {
        s_min[wv] = mn;
        s_max[wv] = mx;
        s_zero[wv] = zc;
    }
    __syncthreads();
    uint32_t zeros = 0;
    for (int w = 0; w < kWaves; ++w) {
        mn = s_min[w] < mn ? s_min[w] : mn;
        mx = s_max[w] > mx ? s_max[w] : mx;
        zeros += s_zero[w];
    }
    // ranks among the non-zero keys (<= 0: the target is a zero)
    const int64_t t0 = (int64_t)bp.klo - zeros, t1 = (int64_t)bp.khi - zeros;
    const U range = ((int64_t)zeros < m) ? (U)(mx - mn) : (U)0;
    const int B = range ? NB - (int)(sizeof(U) == 8 ? __builtin_clzll((uint64_t)range)
                                                    : __builtin_clz((uint32_t)range))
                        : 0;
    if (tid < 2) {
        s_pre[tid] = 0;
        s_k[tid] = tid ? t1 : t0;
    }
    __syncthreads();
    int known = 0;
    for (int pass = 0; pass < 2 && known < B; ++pass) {
        const int width = (B - known) < kDigit ? (B - known) : kDigit;
        const int shift = B - known - width;
        const U dmask = (U)((1u << width) - 1u);
        for (int b = tid; b < 2 * kSelBins; b += kSelBlock) (&hist[0][0])[b] = 0;
        __syncthreads();
        const U p0 = s_pre[0], p1 = s_pre[1];
        const bool h0 = s_k[0] > 0, h1 = s_k[1] > 0;
        for (int i = tid; i < m; i += kSelBlock) {
            const U key = keys[i];
            if (key == 0) continue;
            const U d = key - mn;
            const uint32_t dg = (uint32_t)((d >> shift) & dmask);
            const bool same0 = known == 0 || (d >> (B - known)) == (p0 >> (B - known));
            const bool same1 = known == 0 || (d >> (B - known)) == (p1 >> (B - known));
            if (h0 && same0) atomicAdd(&hist[0][dg], 1u);
            if (h1 && same1) atomicAdd(&hist[1][dg], 1u);
        }
        __syncthreads();
        const int64_t kk0 = s_k[0], kk1 = s_k[1];
        const uint32_t a0 = hist[0][2 * tid], a1 = hist[0][2 * tid + 1];
        const uint32_t b0 = hist[1][2 * tid], b1 = hist[1][2 * tid + 1];
        uint32_t ia = a0 + a1, ib = b0 + b1;
        for (int o = 1; o < kWave; o <<= 1) {
            const uint32_t va = __shfl_up(ia, o, kWave), vb = __shfl_up(ib, o, kWave);
            if (lane >= o) {
                ia += va;
                ib += vb;
            }
        }
        if (lane == kWave - 1) {
            wtot[0][wv] = ia;
            wtot[1][wv] = ib;
        }
        __syncthreads();
        uint32_t ba = 0, bb = 0;
        for (int q = 0; q < wv; ++q) {
            ba += wtot[0][q];
            bb += wtot[1][q];
        }
        const int64_t ea = (int64_t)ba + ia - (a0 + a1), eb = (int64_t)bb + ib - (b0 + b1);
        if (ea < kk0 && kk0 <= ea + a0) {
            s_pre[0] = p0 | ((U)(2 * tid) << shift);
            s_k[0] = kk0 - ea;
        } else if (ea + a0 < kk0 && kk0 <= ea + a0 + a1) {
            s_pre[0] = p0 | ((U)(2 * tid + 1) << shift);
            s_k[0] = kk0 - ea - a0;
        }
        if (eb < kk1 && kk1 <= eb + b0) {
            s_pre[1] = p1 | ((U)(2 * tid) << shift);
            s_k[1] = kk1 - eb;
        } else if (eb + b0 < kk1 && kk1 <= eb + b0 + b1) {
            s_pre[1] = p1 | ((U)(2 * tid + 1) << shift);
            s_k[1] = kk1 - eb - b0;
        }
        __syncthreads();
        known += width;
    }
    if (tid == 0) {
        const U rest = (B - known) > 0 ? (U)(((U)1 << (B - known)) - 1) : (U)0;
        SelRowState st{};
        st.lo = (bp.klo <= 1 || t0 <= 0) ? (uint64_t)0 : (uint64_t)(U)(mn + s_pre[0]);
        st.hi = bp.khi >= m ? (uint64_t)(U)~(U)0
                : t1 <= 0   ? (uint64_t)0
                            : (uint64_t)(U)(mn + (s_pre[1] | rest));
        state[row] = st;
    }
}

// One pass over a chunk of a row: counts below lo / zeros, and stages the non-zero entries
// inside [lo, hi] in LDS. Each wave reads its own kCountSlice entries of the chunk, all of
// its loads issued before the first is counted, and stages its inside entries in its own
// kCountStage-entry slice of LDS, so the staging needs no atomics: the slot of a lane is the
// wave's running count plus the lanes below it in the ballot. The slice holds several times
// the expected inside count of a wave's share at any quantile (the bracket covers at most
// 12 sample standard deviations + 4 ranks, under 7 % of the sample); a wave whose inside
// entries do not fit flags the row (kSelOverflow in its claim), whose selection then runs
// over the whole row.
constexpr uint32_t kSelOverflow = 0x80000000u;

// Adds n to a row's inside count and takes `slots` candidate slots with it: one 64-bit
// atomic on the adjacent words (nin, claim), nin in the low half (a row's count stays below
// 2^32, so it never carries into claim). Returns claim before the add, kSelOverflow
// included if an earlier counter flagged the row.
__device__ __forceinline__ uint32_t sel_count_and_claim(SelRowState* s, uint32_t n, uint32_t slots)
{
    static_assert(offsetof(SelRowState, nin) % 8 == 0 &&
                      offsetof(SelRowState, claim) == offsetof(SelRowState, nin) + 4,
                  "nin and claim form one aligned 64-bit word");
    const unsigned long long old = atomicAdd(reinterpret_cast<unsigned long long*>(&s->nin),
                                             ((unsigned long long)slots << 32) | n);
    return (uint32_t)(old >> 32);
}
constexpr int kCountSlice = kCountChunk / (kCountBlock / kWave);  // entries read per wave
constexpr int kCountStage = 512;                                  // entries staged per wave

template <typename T>
__global__ __launch_bounds__(kCountBlock) void bracket_count_kernel(const T* __restrict__ x, int64_t n,
                                                                    int chunks,
                                                                    SelRowState* __restrict__ state,
                                                                    T* __restrict__ cand, int64_t cap)
{
    typedef typename SelectBits<T>::U U;
    constexpr int kWaves = kCountBlock / kWave;
    __shared__ U buf[kWaves * kCountStage];
    __shared__ uint32_t s_cnt[3][kWaves];
    __shared__ uint32_t s_base;
    const int row = blockIdx.x / chunks, ch = blockIdx.x % chunks;
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wv = tid / kWave;
    const int64_t beg = (int64_t)ch * kCountChunk;
    const int len = (int)((n - beg) < kCountChunk ? (n - beg) : kCountChunk);
    const U lo = (U)state[row].lo, hi = (U)state[row].hi;
    const U lo1 = lo ? lo : (U)1;  // inside: non-zero and >= lo
    const bool any_in = hi >= lo1;
    const U span = any_in ? (U)(hi - lo1) : (U)0;
    const bool collect = any_in && lo != hi;
    const T* r = x + (int64_t)row * n + beg;
    const uint64_t below_me = lane ? (~0ull >> (64 - lane)) : 0ull;
    U* slice = buf + wv * kCountStage;
    uint32_t cnt_lo = 0;  // entries < lo1: below lo (lo > 0), or the zeros (lo = 0)
    uint32_t nin = 0;     // wave-uniform
    bool ovf = false;     // wave-uniform: an inside entry was not staged
    auto visit = [&](U b, bool valid) {
        cnt_lo += (valid && b < lo1) ? 1u : 0u;
        const bool in = valid && any_in && (U)(b - lo1) <= span;
        const uint64_t mask = __ballot(in);
        const uint32_t c = (uint32_t)__popcll(mask);
        if (collect && c) {
            if (!ovf && nin + c <= (uint32_t)kCountStage) {
                if (in) slice[nin + (uint32_t)__popcll(mask & below_me)] = b;
            } else {
                ovf = true;
            }
        }
        nin += c;
    };
    constexpr int V = 16 / (int)sizeof(T);
    constexpr int NV = kCountSlice / (V * kWave);  // vectors per lane of a whole slice
    typedef U Vec __attribute__((ext_vector_type(V)));
    const int wbeg = wv * kCountSlice;
    const int wlen = len - wbeg < 0 ? 0 : (len - wbeg < kCountSlice ? len - wbeg : kCountSlice);
    const U* wr = reinterpret_cast<const U*>(r) + wbeg;
    if ((((uintptr_t)wr) & 15) == 0 && wlen == kCountSlice) {
        const Vec* rv = reinterpret_cast<const Vec*>(wr);
        Vec v[NV];
#pragma unroll
        for (int j = 0; j < NV; ++j) v[j] = __builtin_nontemporal_load(rv + j * kWave + lane);
#pragma unroll
        for (int j = 0; j < NV; ++j)
#pragma unroll
            for (int e = 0; e < V; ++e) visit(v[j][e], true);
    } else {
        for (int i0 = 0; i0 < wlen; i0 += kWave) {
            const int i = i0 + lane;
            const bool valid = i < wlen;
            visit(valid ? wr[i] : (U)0, valid);
        }
    }
    for (int o = kWave / 2; o > 0; o >>= 1) cnt_lo += __shfl_xor(cnt_lo, o, kWave);
    if (lane == 0) {
        s_cnt[0][wv] = cnt_lo;
        s_cnt[1][wv] = nin;
        s_cnt[2][wv] = ovf ? 1u : 0u;
    }
    __syncthreads();
    bool any_ovf = false;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) any_ovf |= s_cnt[2][w] != 0u;
    if (tid == 0) {
        uint32_t b = 0, c = 0;
        for (int w = 0; w < kWaves; ++w) {
            b += s_cnt[0][w];
            c += s_cnt[1][w];
        }
        if (b) atomicAdd(lo ? &state[row].below : &state[row].zin, b);
        const uint32_t take = (collect && !any_ovf) ? c : 0u;
        const uint32_t base = c ? sel_count_and_claim(&state[row], c, take) : 0u;
        if (collect && any_ovf) atomicOr(&state[row].claim, kSelOverflow);
        s_base = take ? base : 0u;
    }
    __syncthreads();
    if (collect && !any_ovf) {
        uint32_t off = s_base;
        for (int w = 0; w < wv; ++w) off += s_cnt[1][w];
        const uint32_t cnt = s_cnt[1][wv];
        T* out = cand + (int64_t)row * cap;
        for (uint32_t i = lane; i < cnt; i += kWave)
            if ((int64_t)off + i < cap) out[off + i] = __builtin_bit_cast(T, slice[i]);
    }
}

// True when the counting pass's results answer the row (bracket_finish_kernel then reads
// only the counts and the candidates); false when the sample missed rank k or the
// candidates did not all fit, and the selection must see the whole row.
__host__ __device__ inline bool bracket_row_resolved(const SelRowState& s, int64_t k, int64_t cap)
{
    const int64_t below = s.below, zin = s.zin, nin = s.nin;
    if (!(below < k && k <= below + zin + nin)) return false;
    if (k - below <= zin || s.lo == s.hi) return true;
    // (the top bit of claim: a counting pass could not stage every candidate)
    return nin <= cap && !(s.claim & kSelOverflow);
}

template <typename T>
__global__ __launch_bounds__(kSelBlock) void bracket_finish_kernel(const T* __restrict__ x, int64_t n,
                                                                   int64_t k,
                                                                   const SelRowState* __restrict__ state,
                                                                   const T* __restrict__ cand,
                                                                   int64_t cap, double* __restrict__ out)
{
    typedef typename SelectBits<T>::U U;
    const int row = blockIdx.x;
    const SelRowState s = state[row];
    if (bracket_row_resolved(s, k, cap)) {
        const int64_t rk = k - (int64_t)s.below;
        if (rk <= (int64_t)s.zin) {
            if (threadIdx.x == 0) out[row] = 0.0;
        } else if ((U)s.lo == (U)s.hi) {
            if (threadIdx.x == 0) out[row] = (double)__builtin_bit_cast(T, (U)s.lo);
        } else {
            select_kth_row<T>(cand + (int64_t)row * cap, s.nin, rk - (int64_t)s.zin, out + row, true);
        }
        return;
    }
    select_kth_row<T>(x + (int64_t)row * n, n, k, out + row);
}

template <typename T>
inline void bracket_ws_split(void* ws, int rows, SelRowState** state, T** cand)
{
    const size_t head = ((size_t)rows * sizeof(SelRowState) + 255) & ~(size_t)255;
    *state = static_cast<SelRowState*>(ws);
    *cand = reinterpret_cast<T*>(static_cast<char*>(ws) + head);
}

// k-th smallest (1-based) of each of `rows` rows of n non-negative values into out[rows].
// ws: device workspace of bracket_workspace_bytes(rows, bracket_plan(n, k, sizeof(T)),
// sizeof(T)) bytes, or null for the whole-row form.
template <typename T>
hipError_t launch_select_rows(const T* x, int rows, int64_t n, int64_t k, void* ws, double* out,
                              hipStream_t st)
{
    if (rows <= 0) return hipSuccess;
    const BracketPlan bp = bracket_plan(n, k, (int)sizeof(T));
    if (!bp.use || !ws) {
        select_kth_kernel<T><<<rows, kSelBlock, 0, st>>>(x, n, k, out);
        return hipGetLastError();
    }
    if ((int64_t)rows * bp.chunks >= ((int64_t)1 << 31)) return hipErrorInvalidValue;
    SelRowState* state;
    T* cand;
    bracket_ws_split<T>(ws, rows, &state, &cand);
    bracket_sample_kernel<T><<<rows, kSelBlock, 0, st>>>(x, n, bp, state);
    bracket_count_kernel<T><<<(unsigned)(rows * bp.chunks), kCountBlock, 0, st>>>(x, n, bp.chunks,
                                                                                   state, cand, bp.cap);
    bracket_finish_kernel<T><<<rows, kSelBlock, 0, st>>>(x, n, k, state, cand, bp.cap, out);
    return hipGetLastError();
}

// The first and last launches of launch_select_rows, for a caller whose own pass over the
// rows does the counting (bracket_plan(n, k, sizeof(T)).use must hold and ws be set).
template <typename T>
hipError_t launch_bracket_sample_rows(const T* x, int rows, int64_t n, int64_t k, void* ws,
                                      hipStream_t st)
{
    const BracketPlan bp = bracket_plan(n, k, (int)sizeof(T));
    if (!bp.use || !ws || rows <= 0) return hipErrorInvalidValue;
    SelRowState* state;
    T* cand;
    bracket_ws_split<T>(ws, rows, &state, &cand);
    bracket_sample_kernel<T><<<rows, kSelBlock, 0, st>>>(x, n, bp, state);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_bracket_finish_rows(const T* x, int rows, int64_t n, int64_t k, void* ws,
                                      double* out, hipStream_t st)
{
    const BracketPlan bp = bracket_plan(n, k, (int)sizeof(T));
    if (!bp.use || !ws || rows <= 0) return hipErrorInvalidValue;
    SelRowState* state;
    T* cand;
    bracket_ws_split<T>(ws, rows, &state, &cand);
    bracket_finish_kernel<T><<<rows, kSelBlock, 0, st>>>(x, n, k, state, cand, bp.cap, out);
    return hipGetLastError();
}

}  // namespace hjm
