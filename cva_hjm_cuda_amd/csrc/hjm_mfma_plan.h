// hjm_mfma_plan.h -- host construction of the 50-step block maps for the matrix-core
// path kernels (FP32 as bf16 parts, FP64 as they are) and of the valuation's fragments.
//
// One block of S = 50 Euler steps is affine in the curve f, the scaled shocks
// s_{t,k} = sqrt(dt) phi_k(t) and 1:
//   f_out  = T^S f_in + sum_t T^{S-1-t} (m dt + V s_t)
// with T = I + (dt/dtau) Dfwd (backward difference on the last tenor). Two maps:
//   reference semantics (102 rows, 7 M-tiles of 16): rows 51..101 hold the block's
//     discount-integral increments dA = sum_{r=1}^{S} f_r
//       = (sum_{r=1}^{S} T^r) f_in + sum_t (sum_{q=0}^{S-1-t} T^q)(m dt + V s_t);
//   portfolio semantics (52 rows, 4 M-tiles): row 51 holds the short rate's left-point sum
//       sum_{r=0}^{S-1} f_r[0] = (sum_{r=0}^{S-1} T^r f_in)[0]
//                               + sum_t (sum_{q=0}^{S-2-t} T^q (m dt + V s_t))[0].
// The kernel computes W x [f_in; phi; 1] with v_mfma_f32_16x16x32_bf16. K has 224 entries
// in 7 chunks of 32; in chunk c lane group g = lane / 16 supplies slots i = 0..7, the
// K index being 32 c + 8 g + i:
//   chunks 0, 1: curve row R = 16 (2c + i/4) + 4 g + i%4 (zero column where R > 50): the
//     previous block's accumulator registers 8c + i of the lane, unchanged;
//   chunks 2..6 (s = c - 2): value v = 8 s + i of lane group g, which generates the steps
//     t = 4 q + g, q = 0..12: v = 3 q + k is factor k of step t (t < 50); at t = 50
//     (g = 2, q = 12) factor 0 is the constant 1 carrying the drift; other slots are zero.
// W (FP64) is split into three bf16 parts W = Wh + Wm + Wl (round to nearest); the kernel
// splits its FP32 operand x exactly (truncation: xh + xm + xl = x) and accumulates the six
// products Wh xh + Wh xm + Wm xh + Wh xl + Wl xh + Wm xm in FP32, which leaves an error of
// the order of FP32 rounding.
// Fragment [m][c][part][lane][i] = part of W[16 m + lane % 16][32 c + 8 (lane / 16) + i].
#pragma once
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "hjm_common.h"

namespace hjm {

constexpr int kMfmaStride = 50;   // steps per block handled by the matrix-core kernels
constexpr int kMfmaMTiles = 7;    // reference map: 102 output rows -> 7 x 16
constexpr int kMfmaPfMTiles = 4;  // portfolio map: 52 output rows -> 4 x 16
constexpr int kMfmaQuads = 13;    // steps 4 q + g, q = 0..12, per lane group
constexpr int kMfmaChunks = 7;    // 2 curve chunks + 5 shock chunks of 32 K entries
constexpr int kMfmaK = 32 * kMfmaChunks;  // 224
constexpr int kMfmaParts = 3;     // bf16 parts of W
// bf16 elements of one kernel's fragments
constexpr int kMfmaFragElems = kMfmaMTiles * kMfmaChunks * kMfmaParts * kWave * 8;
constexpr int kMfmaPfFragElems = kMfmaPfMTiles * kMfmaChunks * kMfmaParts * kWave * 8;
// LDS bytes of the first mt M-tiles' fragments (the reference kernel stages only its live
// tiles)
constexpr int ref_mfma_lds_bytes(int mt) { return mt * kMfmaChunks * kMfmaParts * kWave * 8 * 2; }

enum class MfmaMap { kReference, kPortfolio };

// Live M-tiles of the reference map for a run of ndates dates: the curve's four and the
// integral rows 51 + j up to j = ndates (the last slot recorded).
inline int ref_mfma_tiles(int ndates)
{
    const int last = kTenors + (ndates > 1 ? ndates : 1);
    return last / 16 + 1 < 4 ? 4 : last / 16 + 1;
}

// Structural zeros of the maps' curve columns. T couples tenor j only to j + 1 (and 50 to
// 49), so T^n and the integral rows' sums of powers vanish in columns R < min(j, 49) of
// row j; integral row 51 + j follows tenor j; the portfolio's row 51 (the short rate's
// sum) spans every tenor. The kernels skip a product whose curve inputs all lie below the
// first input of every row of its M-tile.
__host__ __device__ constexpr int map_row_first_input(int row, bool pf)
{
    return row < kTenors ? (row < 49 ? row : 49) : pf ? 0 : (row - kTenors < 49 ? row - kTenors : 49);
}
__host__ __device__ constexpr int tile_first_input(int m, bool pf)
{
    const int rows = pf ? kTenors + 1 : 2 * kTenors;
    int lo = kTenors;
    for (int r = 16 * m; r < 16 * m + 16 && r < rows; ++r)
        lo = map_row_first_input(r, pf) < lo ? map_row_first_input(r, pf) : lo;
    return lo;
}
// FP32 kernels: curve chunk c (rows 32 c .. 32 c + 31) of M-tile m is all zero
__host__ __device__ constexpr bool mfma_chunk_zero(int m, int c, bool pf)
{
    return c < 2 && 32 * c + 31 < tile_first_input(m, pf);
}

// What K index k of the block map multiplies: a curve row, a shock (step, factor), the
// constant 1, or nothing.
struct MfmaColumn {
    enum Kind { kZero, kCurve, kShock, kConst } kind;
    int row;     // kCurve
    int step;    // kShock
    int factor;  // kShock
};

inline MfmaColumn mfma_column(int k)
{
    const int c = k / 32, g = (k % 32) / 8, i = k % 8;
    if (c < 2) {
        const int R = 16 * (2 * c + i / 4) + 4 * g + i % 4;
        if (R < kTenors) return {MfmaColumn::kCurve, R, 0, 0};
        return {MfmaColumn::kZero, 0, 0, 0};
    }
    const int v = 8 * (c - 2) + i;
    if (v >= 3 * kMfmaQuads) return {MfmaColumn::kZero, 0, 0, 0};
    const int q = v / 3, f = v % 3, t = 4 * q + g;
    if (t < kMfmaStride) return {MfmaColumn::kShock, 0, t, f};
    if (t == kMfmaStride && f == 0) return {MfmaColumn::kConst, 0, 0, 0};
    return {MfmaColumn::kZero, 0, 0, 0};
}

// Dense FP64 block map [rows][224]: rows = 102 (reference) or 52 (portfolio).
inline std::vector<double> mfma_block_map_dense(const double* drift, const double* vols, double dt,
                                                double dtau, MfmaMap map)
{
    const bool pf = map == MfmaMap::kPortfolio;
    const int n = kTenors, S = kMfmaStride;
    const int rows = pf ? n + 1 : 2 * n;
    const double c = dt / dtau, sq = sqrt(dt);
    auto matmul = [n](const std::vector<double>& A, const std::vector<double>& B) {
        std::vector<double> C((size_t)n * n, 0.0);
        for (int i = 0; i < n; ++i)
            for (int k = 0; k < n; ++k) {
                const double a = A[(size_t)i * n + k];
                if (a == 0.0) continue;
                for (int j = 0; j < n; ++j) C[(size_t)i * n + j] += a * B[(size_t)k * n + j];
            }
        return C;
    };
    std::vector<double> T((size_t)n * n, 0.0);
    for (int j = 0; j < n - 1; ++j) {
        T[(size_t)j * n + j] = 1.0 - c;
        T[(size_t)j * n + j + 1] = c;
    }
    T[(size_t)(n - 1) * n + n - 1] = 1.0 + c;
    T[(size_t)(n - 1) * n + n - 2] = -c;
    // Tp[s] = T^s, Cs[s] = sum_{q=0}^{s} T^q
    std::vector<std::vector<double>> Tp(S + 1), Cs(S + 1);
    Tp[0].assign((size_t)n * n, 0.0);
    for (int j = 0; j < n; ++j) Tp[0][(size_t)j * n + j] = 1.0;
    Cs[0] = Tp[0];
    for (int s = 1; s <= S; ++s) {
        Tp[s] = matmul(T, Tp[s - 1]);
        Cs[s] = Cs[s - 1];
        for (size_t i = 0; i < Cs[s].size(); ++i) Cs[s][i] += Tp[s][i];
    }
    auto mv = [n](const std::vector<double>& M, const double* v, int j) {
        double s = 0.0;
        for (int i = 0; i < n; ++i) s += M[(size_t)j * n + i] * v[i];
        return s;
    };
    std::vector<double> mdt(n);
    for (int i = 0; i < n; ++i) mdt[i] = drift[i] * dt;
    std::vector<double> W((size_t)rows * kMfmaK, 0.0);
    for (int k = 0; k < kMfmaK; ++k) {
        const MfmaColumn col = mfma_column(k);
        auto put = [&](int row, double v) { W[(size_t)row * kMfmaK + k] = v; };
        if (col.kind == MfmaColumn::kCurve) {
            const int R = col.row;
            for (int j = 0; j < n; ++j) put(j, Tp[S][(size_t)j * n + R]);
            if (pf) {
                put(n, Cs[S - 1][R]);
            } else {
                for (int j = 0; j < n; ++j) {
                    double sa = 0.0;
                    for (int q = 1; q <= S; ++q) sa += Tp[q][(size_t)j * n + R];
                    put(n + j, sa);
                }
            }
        } else if (col.kind == MfmaColumn::kShock) {
            const int t = col.step;
            std::vector<double> sv(n);
            for (int i = 0; i < n; ++i) sv[i] = sq * vols[col.factor * n + i];
            for (int j = 0; j < n; ++j) put(j, mv(Tp[S - 1 - t], sv.data(), j));
            if (pf) {
                if (t <= S - 2) put(n, mv(Cs[S - 2 - t], sv.data(), 0));
            } else {
                for (int j = 0; j < n; ++j) put(n + j, mv(Cs[S - 1 - t], sv.data(), j));
            }
        } else if (col.kind == MfmaColumn::kConst) {
            for (int j = 0; j < n; ++j) {
                double sf = 0.0, sa = 0.0;
                for (int tt = 0; tt < S; ++tt) {
                    sf += mv(Tp[S - 1 - tt], mdt.data(), j);
                    if (!pf) sa += mv(Cs[S - 1 - tt], mdt.data(), j);
                    else if (j == 0 && tt <= S - 2) sa += mv(Cs[S - 2 - tt], mdt.data(), 0);
                }
                put(j, sf);
                if (!pf) put(n + j, sa);
                else if (j == 0) put(n, sa);
            }
        }
    }
    return W;
}

// bf16 bits of x rounded to nearest (ties to even); x finite.
inline uint16_t bf16_rne(double x)
{
    const float f = (float)x;
    uint32_t u;
    memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

inline double bf16_value(uint16_t b)
{
    const uint32_t u = (uint32_t)b << 16;
    float f;
    memcpy(&f, &u, 4);
    return (double)f;
}

// Fragments of the three bf16 parts of W in the kernel's A-operand order.
inline void build_mfma_block_map(const double* drift, const double* vols, double dt, double dtau,
                                 std::vector<uint16_t>& frag, MfmaMap map = MfmaMap::kReference)
{
    const bool pf = map == MfmaMap::kPortfolio;
    const int mtiles = pf ? kMfmaPfMTiles : kMfmaMTiles;
    const int rows = pf ? kTenors + 1 : 2 * kTenors;
    const std::vector<double> W = mfma_block_map_dense(drift, vols, dt, dtau, map);
    frag.assign((size_t)mtiles * kMfmaChunks * kMfmaParts * kWave * 8, 0);
    for (int m = 0; m < mtiles; ++m)
        for (int c = 0; c < kMfmaChunks; ++c)
            for (int l = 0; l < kWave; ++l)
                for (int i = 0; i < 8; ++i) {
                    const int row = 16 * m + l % 16, k = 32 * c + 8 * (l / 16) + i;
                    const double w = row < rows ? W[(size_t)row * kMfmaK + k] : 0.0;
                    const uint16_t h = bf16_rne(w);
                    const double r1 = w - bf16_value(h);
                    const uint16_t mm = bf16_rne(r1);
                    const uint16_t lo = bf16_rne(r1 - bf16_value(mm));
                    const uint16_t parts[3] = {h, mm, lo};
                    for (int p = 0; p < kMfmaParts; ++p)
                        frag[((((size_t)m * kMfmaChunks + c) * kMfmaParts + p) * kWave + l) * 8 + i] = parts[p];
                }
}

// The FP64 block map for v_mfma_f64_16x16x4f64 (ref_paths_mfma64_kernel). One instruction
// takes K = 4 entries, one per lane group g = lane / 16, from one register of every lane;
// its accumulator holds row 4 i + g of the 16-row tile in register i of lane group g (not
// 4 g + i as in the FP32 16x16 forms). The kernel issues kMfma64K = 52 instructions per
// M-tile and block:
//   j = 0..12 (curve): register i = j % 4 of the previous block's accumulator tile j / 4,
//     i.e. curve row R = 16 (j / 4) + 4 i + g (zero column where R > 50; tile 3 holds
//     rows 48..50 in register 0 only, hence 13 and not 16);
//   j = 13..51 (shocks): value v = j - 13 of lane group g, factor v % 3 of step
//     4 (v / 3) + g -- the FP32 kernel's shocks, in its order; at step 50 factor 0 is the
//     constant 1 carrying the drift, other entries past step 49 are zero.
// Fragment [m][j][lane] = W[16 m + lane % 16][column (j, lane / 16)], FP64; the reference
// map (ref_paths_mfma64_kernel, up to 7 M-tiles) or the portfolio map (pf_paths_mfma64_kernel,
// 4 M-tiles; the short rate's sum, row 51, is register 0 of lane group 3 of tile 3).
constexpr int kMfma64Curve = 13;
constexpr int kMfma64K = kMfma64Curve + 3 * kMfmaQuads;  // 52
constexpr int kMfma64TileBytes = kMfma64K * kWave * 8;   // 26624
constexpr int kMfma64LdsTiles = 5;  // M-tiles staged in LDS (133 KB); later tiles from HBM
// curve instruction j < 13 of M-tile m is all zero (its rows 16 (j / 4) + 4 (j % 4) + g lie
// below the tile's first input; tiles 0..3 agree for both maps)
__host__ __device__ constexpr bool mfma64_column_zero(int m, int j)
{
    return j < kMfma64Curve && 16 * (j / 4) + 4 * (j % 4) + 3 < tile_first_input(m, false);
}

inline MfmaColumn mfma64_column(int j, int g)
{
    if (j < kMfma64Curve) {
        const int R = 16 * (j / 4) + 4 * (j % 4) + g;
        if (R < kTenors) return {MfmaColumn::kCurve, R, 0, 0};
        return {MfmaColumn::kZero, 0, 0, 0};
    }
    const int v = j - kMfma64Curve, q = v / 3, f = v % 3, t = 4 * q + g;
    if (t < kMfmaStride) return {MfmaColumn::kShock, 0, t, f};
    if (t == kMfmaStride && f == 0) return {MfmaColumn::kConst, 0, 0, 0};
    return {MfmaColumn::kZero, 0, 0, 0};
}

inline void build_mfma64_block_map(const double* drift, const double* vols, double dt, double dtau,
                                   int mtiles, std::vector<double>& frag,
                                   MfmaMap map = MfmaMap::kReference)
{
    const int rows = map == MfmaMap::kPortfolio ? kTenors + 1 : 2 * kTenors;
    const std::vector<double> W = mfma_block_map_dense(drift, vols, dt, dtau, map);
    // the dense map's column of each curve row, shock (step, factor) and the constant
    int curve_k[kTenors] = {}, shock_k[kMfmaStride][3] = {}, const_k = -1;
    for (int k = 0; k < kMfmaK; ++k) {
        const MfmaColumn c = mfma_column(k);
        if (c.kind == MfmaColumn::kCurve) curve_k[c.row] = k;
        else if (c.kind == MfmaColumn::kShock) shock_k[c.step][c.factor] = k;
        else if (c.kind == MfmaColumn::kConst) const_k = k;
    }
    frag.assign((size_t)mtiles * kMfma64K * kWave, 0.0);
    for (int m = 0; m < mtiles; ++m)
        for (int j = 0; j < kMfma64K; ++j)
            for (int l = 0; l < kWave; ++l) {
                const int row = 16 * m + l % 16;
                const MfmaColumn c = mfma64_column(j, l / 16);
                int k = -1;
                if (c.kind == MfmaColumn::kCurve) k = curve_k[c.row];
                else if (c.kind == MfmaColumn::kShock) k = shock_k[c.step][c.factor];
                else if (c.kind == MfmaColumn::kConst) k = const_k;
                if (row < rows && k >= 0)
                    frag[((size_t)m * kMfma64K + j) * kWave + l] = W[(size_t)row * kMfmaK + k];
            }
}

// Matrix-core valuation of netting sets (pf_value_mfma_kernel): per date, M-tiles of 16
// sets x 2 chunks of 32 periods x 3 bf16 parts of -gamma, in the A-operand order
// [date][m][chunk][part][lane][i] = part of -gamma[date][16 m + lane % 16][32 chunk +
// 8 (lane / 16) + i] (0 past the last set or period); apad[date][16 m + j] = a of set
// 16 m + j (0 past the last set); tilek[date][m] = the largest kmax of the tile's sets.
constexpr int kValueFragsPerTile = 2 * kMfmaParts;  // u4 fragments per lane and M-tile

inline int value_mtiles(int nsets) { return (nsets + 15) / 16; }
constexpr int kMaxPfMTiles = (HJM_MAX_SETS + 15) / 16;

inline void build_value_fragments(const std::vector<double>& coef, const std::vector<double>& a,
                                  const std::vector<int>& kmax, int ndates, int nsets, int kpad,
                                  std::vector<uint16_t>& frag, std::vector<float>& apad,
                                  std::vector<int>& tilek)
{
    const int mt = value_mtiles(nsets);
    frag.assign((size_t)ndates * mt * kValueFragsPerTile * kWave * 8, 0);
    apad.assign((size_t)ndates * mt * 16, 0.f);
    tilek.assign((size_t)ndates * mt, 0);
    for (int e = 0; e < ndates; ++e)
        for (int m = 0; m < mt; ++m) {
            for (int j = 0; j < 16; ++j) {
                const int g = 16 * m + j;
                if (g >= nsets) continue;
                apad[((size_t)e * mt + m) * 16 + j] = (float)a[(size_t)e * nsets + g];
                tilek[(size_t)e * mt + m] = std::max(tilek[(size_t)e * mt + m], kmax[(size_t)e * nsets + g]);
            }
            for (int c = 0; c < 2; ++c)
                for (int l = 0; l < kWave; ++l)
                    for (int i = 0; i < 8; ++i) {
                        const int g = 16 * m + l % 16, kk = 32 * c + 8 * (l / 16) + i;
                        const double w = (g < nsets && kk < kpad)
                                             ? -coef[((size_t)e * nsets + g) * kpad + kk]
                                             : 0.0;
                        const uint16_t h = bf16_rne(w);
                        const double r1 = w - bf16_value(h);
                        const uint16_t mm = bf16_rne(r1);
                        const uint16_t lo = bf16_rne(r1 - bf16_value(mm));
                        const uint16_t parts[3] = {h, mm, lo};
                        for (int q = 0; q < kMfmaParts; ++q)
                            frag[(((((size_t)e * mt + m) * 2 + c) * kMfmaParts + q) * kWave + l) * 8 + i] =
                                parts[q];
                    }
        }
}

// The portfolio adjoint's second product dP = (-Gamma)^T u on the matrix cores
// (pf_adjoint_seed_mfma_kernel): per date, 4 period tiles x ceil(mtiles / 2) K chunks of 32
// sets x 3 bf16 parts, [date][pm][c][part][lane][i] = part of -gamma[date][set][16 pm +
// lane % 16] with set = 16 (2c + i/4) + 4 (lane/16) + i%4, the order in which the first
// product's accumulators hold u (0 past the last set or period).
inline void build_adjoint_fragments(const std::vector<double>& coef, int ndates, int nsets, int kpad,
                                    std::vector<uint16_t>& frag)
{
    const int nch = (value_mtiles(nsets) + 1) / 2;
    frag.assign((size_t)ndates * 4 * nch * kMfmaParts * kWave * 8, 0);
    for (int e = 0; e < ndates; ++e)
        for (int pm = 0; pm < 4; ++pm)
            for (int c = 0; c < nch; ++c)
                for (int l = 0; l < kWave; ++l)
                    for (int i = 0; i < 8; ++i) {
                        const int set = 16 * (2 * c + i / 4) + 4 * (l / 16) + i % 4;
                        const int kk = 16 * pm + l % 16;
                        const double w = (set < nsets && kk < kpad)
                                             ? -coef[((size_t)e * nsets + set) * kpad + kk]
                                             : 0.0;
                        const uint16_t h = bf16_rne(w);
                        const double r1 = w - bf16_value(h);
                        const uint16_t mm = bf16_rne(r1);
                        const uint16_t lo = bf16_rne(r1 - bf16_value(mm));
                        const uint16_t parts[3] = {h, mm, lo};
                        for (int q = 0; q < kMfmaParts; ++q)
                            frag[(((((size_t)e * 4 + pm) * nch + c) * kMfmaParts + q) * kWave + l) * 8 + i] =
                                parts[q];
                    }
}

}  // namespace hjm
