// hjm_mfma_plan.h -- host construction of the 50-step block maps for the matrix-core
// FP32 path kernels.
//
// One block of `stride` Euler steps is affine in the curve f, the scaled shocks
// s_{t,k} = sqrt(dt) phi_k(t) and 1:
//   f_out  = T^S f_in + sum_t T^{S-1-t} (m dt + V s_t)
// with T = I + (dt/dtau) Dfwd (backward difference on the last tenor). Two maps:
//   reference semantics: rows 51..101 hold the block's discount-integral increment
//     dA = sum_{r=1}^{S} f_r = (sum_{r=1}^{S} T^r) f_in + sum_t (sum_{q=0}^{S-1-t} T^q)(m dt + V s_t)
//     (7 M-tiles of 16 rows);
//   portfolio semantics: row 51 holds the short rate's left-point sum
//     sum_{r=0}^{S-1} f_r[0] = (sum_{r=0}^{S-1} T^r f_in)[0] + sum_t (sum_{q=0}^{S-2-t} T^q (m dt + V s_t))[0]
//     (4 M-tiles; rows 52..63 are zero).
// The kernel computes W x [f_in; phi (3 per step); 1] with MFMA 16x16x4 f32 tiles. K is
// laid out in chunks of 4 so that each lane group g = lane/16 of a chunk is one K index:
//   chunk kc < 16      : the previous block's output registers, kc = 4 m + r holds rows
//                        16 m + 4 g + r (zero columns where that row is not a curve row)
//   chunk 16 + 3 q + k : factor k of step t = 4 q + g (t < S); step S (q = 12, g = 2) k = 0
//                        is the constant 1 carrying the drift; other padding is zero.
// Wfrag[m][kc][lane] = W[16 m + lane % 16][K(kc, lane / 16)] is the A-operand fragment.
#pragma once
#include <vector>

#include "hjm_common.h"

namespace hjm {

constexpr int kMfmaStride = 50;   // steps per block handled by the matrix-core kernels
constexpr int kMfmaMTiles = 7;    // reference map: 102 output rows -> 7 x 16
constexpr int kMfmaPfMTiles = 4;  // portfolio map: 52 output rows -> 4 x 16
constexpr int kMfmaFChunks = 16;  // previous block's rows 0..63 as B operand
constexpr int kMfmaQuads = 13;    // ceil(50 / 4) step quads
constexpr int kMfmaKChunks = kMfmaFChunks + 3 * kMfmaQuads;  // 55
constexpr int kMfmaFragFloats = kMfmaMTiles * kMfmaKChunks * kWave;
constexpr int kMfmaPfFragFloats = kMfmaPfMTiles * kMfmaKChunks * kWave;

enum class MfmaMap { kReference, kPortfolio };

inline void build_mfma_block_map(const double* drift, const double* vols, double dt, double dtau,
                                 std::vector<float>& wfrag, MfmaMap map = MfmaMap::kReference)
{
    const bool pf = map == MfmaMap::kPortfolio;
    const int mtiles = pf ? kMfmaPfMTiles : kMfmaMTiles;
    const int n = kTenors, S = kMfmaStride;
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
    const int K = 4 * kMfmaKChunks;
    std::vector<double> W((size_t)mtiles * 16 * K, 0.0);
    auto wref = [&](int row, int kidx) -> double& { return W[(size_t)row * K + kidx]; };
    auto mv = [n](const std::vector<double>& M, const double* v, int j) {
        double s = 0.0;
        for (int i = 0; i < n; ++i) s += M[(size_t)j * n + i] * v[i];
        return s;
    };
    // row sums of the integral part: reference row n + j, portfolio row n (tenor 0 only)
    auto integral_cols = [&](int kidx, const std::vector<double>& M, const double* v) {
        if (pf) {
            wref(n, kidx) = v ? mv(M, v, 0) : 0.0;
        } else {
            for (int j = 0; j < n; ++j) wref(n + j, kidx) = v ? mv(M, v, j) : 0.0;
        }
    };
    std::vector<double> mdt(n);
    for (int i = 0; i < n; ++i) mdt[i] = drift[i] * dt;
    for (int kc = 0; kc < kMfmaKChunks; ++kc) {
        for (int g = 0; g < 4; ++g) {
            const int kidx = 4 * kc + g;
            if (kc < kMfmaFChunks) {
                const int m = kc / 4, r = kc % 4, R = 16 * m + 4 * g + r;
                if (R >= n) continue;  // rows 51..63 of the previous output are not curve rows
                for (int j = 0; j < n; ++j) wref(j, kidx) = Tp[S][(size_t)j * n + R];
                if (pf) {
                    wref(n, kidx) = Cs[S - 1][(size_t)0 * n + R];
                } else {
                    for (int j = 0; j < n; ++j) {
                        double sa = 0.0;
                        for (int q = 1; q <= S; ++q) sa += Tp[q][(size_t)j * n + R];
                        wref(n + j, kidx) = sa;
                    }
                }
            } else {
                const int q = (kc - kMfmaFChunks) / 3, k = (kc - kMfmaFChunks) % 3;
                const int t = 4 * q + g;
                if (t < S) {
                    std::vector<double> sv(n);
                    for (int i = 0; i < n; ++i) sv[i] = sq * vols[k * n + i];
                    for (int j = 0; j < n; ++j) wref(j, kidx) = mv(Tp[S - 1 - t], sv.data(), j);
                    if (pf) {
                        if (t <= S - 2) integral_cols(kidx, Cs[S - 2 - t], sv.data());
                    } else {
                        integral_cols(kidx, Cs[S - 1 - t], sv.data());
                    }
                } else if (t == S && k == 0) {
                    for (int j = 0; j < n; ++j) {
                        double sf = 0.0, sa = 0.0;
                        for (int tt = 0; tt < S; ++tt) {
                            sf += mv(Tp[S - 1 - tt], mdt.data(), j);
                            if (!pf) sa += mv(Cs[S - 1 - tt], mdt.data(), j);
                            else if (j == 0 && tt <= S - 2) sa += mv(Cs[S - 2 - tt], mdt.data(), 0);
                        }
                        wref(j, kidx) = sf;
                        if (!pf) wref(n + j, kidx) = sa;
                        else if (j == 0) wref(n, kidx) = sa;
                    }
                }
            }
        }
    }
    wfrag.assign((size_t)mtiles * kMfmaKChunks * kWave, 0.f);
    for (int m = 0; m < mtiles; ++m)
        for (int kc = 0; kc < kMfmaKChunks; ++kc)
            for (int l = 0; l < kWave; ++l) {
                const int row = 16 * m + (l % 16), kidx = 4 * kc + l / 16;
                wfrag[((size_t)m * kMfmaKChunks + kc) * kWave + l] = (float)W[(size_t)row * K + kidx];
            }
}

}  // namespace hjm
