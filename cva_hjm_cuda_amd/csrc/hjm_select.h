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

template <typename T>
__device__ void select_kth_row(const T* __restrict__ row, int64_t n, int64_t k, double* out)
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
