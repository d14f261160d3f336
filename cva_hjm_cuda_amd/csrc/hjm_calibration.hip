// hjm_calibration.hip -- batched PCA calibration of the HJM loadings from historical
// forward curves (the step upstream of the path, reference roadmap doc/TODO:16), gfx950.
//
// For each rolling window of `window` consecutive curves (51 tenors each): daily
// differences d_t = f_{t+1} - f_t, their sample covariance annualised by `obs_per_year`,
// its full eigendecomposition, and the loadings sigma_k = sqrt(lambda_k) v_k of the
// leading factors, optionally smoothed by least-squares polynomials in the tenor. The HJM
// drift of the loadings is the library's cumulative Simpson rule (host, hjm_hjm_drift).
//
// One workgroup (256 threads) per window:
//   1. covariance: the window's differences are staged through LDS in chunks of 32 days;
//      column means first, then centred cross products (1326 entries of the upper
//      triangle spread over the threads), FP64;
//   2. cyclic Jacobi in LDS (52 x 52 with a zero pad row/column): round-robin tournament
//      ordering, 26 disjoint rotations per round applied as one row pass and one column
//      pass, sweeps until the off-diagonal mass is below 1e-28 of the total (at most 30);
//   3. eigenvalues sorted descending, eigenvector signs fixed (sum of components > 0,
//      else the first component above 1e-9 in magnitude positive), loadings and fits.
#include <math.h>
#include <string.h>

#include <vector>

#include "hjm_host.h"

using namespace hjm;
using namespace hjm_host;

namespace {

constexpr int kN = kTenors;        // 51
constexpr int kNP = kTenors + 1;   // 52: even size for the tournament
constexpr int kCalBlock = 256;
constexpr int kChunk = 32;         // days per LDS chunk
constexpr int kMaxSweeps = 30;
constexpr int kMaxFactors = 8;
constexpr int kMaxDegree = 5;

struct CalibParams {
    int64_t ncurves;
    int64_t window;       // curves per window (window - 1 differences)
    int64_t step;         // curves between window starts
    int32_t nwindows;
    int32_t nfactors;
    double obs_per_year;
    double dtau;
    int32_t degree[kMaxFactors];  // < 0: raw PCA loading
};

// Position of player i in round r of the round-robin tournament on kNP players: player 0
// stays, the others rotate; pairs are (pos 2p, pos 2p + 1) after the circle mapping.
__device__ __forceinline__ void tournament_pair(int r, int p, int& a, int& b)
{
    // circle method: slots 0..kNP-1, slot 0 fixed; slot s > 0 holds player 1 + (s - 1 + r) % (kNP - 1)
    auto player = [r](int s) { return s == 0 ? 0 : 1 + (s - 1 + r) % (kNP - 1); };
    const int s0 = p, s1 = kNP - 1 - p;
    a = player(s0);
    b = player(s1);
    if (a > b) {
        const int t = a;
        a = b;
        b = t;
    }
}

__global__ __launch_bounds__(kCalBlock) void calib_pca_kernel(const CalibParams prm,
                                                             const double* __restrict__ curves,
                                                             double* __restrict__ evals_out,
                                                             double* __restrict__ loadings_out,
                                                             double* __restrict__ raw_out)
{
    __shared__ double A[kNP][kNP + 1];
    __shared__ double V[kNP][kNP + 1];
    __shared__ double chunk[kChunk][kN];
    __shared__ double mean[kN];
    __shared__ double cs[kNP / 2], sn[kNP / 2];
    __shared__ double red[kCalBlock / kWave], red2[kCalBlock / kWave];
    __shared__ int order[kNP];

    const int tid = threadIdx.x;
    const int w = blockIdx.x;
    const double* base = curves + (int64_t)w * prm.step * kN;
    const int64_t nd = prm.window - 1;

    // ---- 1. column means of the differences
    if (tid < kN) {
        double s = 0.0;
        for (int64_t t = 0; t < nd; ++t) s += base[(t + 1) * kN + tid] - base[t * kN + tid];
        mean[tid] = s / (double)nd;
    }
    __syncthreads();
    // ---- centred cross products; thread owns entries e = tid, tid + 256, ... of the
    // upper triangle (i <= j), 1326 in all
    constexpr int kPairs = kN * (kN + 1) / 2;
    constexpr int kPerThread = (kPairs + kCalBlock - 1) / kCalBlock;  // 6
    int pi[kPerThread], pj[kPerThread];
    double acc[kPerThread];
#pragma unroll
    for (int u = 0; u < kPerThread; ++u) {
        int e = tid + u * kCalBlock, i = 0;
        acc[u] = 0.0;
        if (e >= kPairs) {
            pi[u] = -1;
            pj[u] = -1;
            continue;
        }
        while (e >= kN - i) {
            e -= kN - i;
            ++i;
        }
        pi[u] = i;
        pj[u] = i + e;
    }
    for (int64_t t0 = 0; t0 < nd; t0 += kChunk) {
        const int len = (int)((nd - t0) < kChunk ? (nd - t0) : kChunk);
        for (int x = tid; x < len * kN; x += kCalBlock) {
            const int t = x / kN, j = x % kN;
            const int64_t tt = t0 + t;
            chunk[t][j] = (base[(tt + 1) * kN + j] - base[tt * kN + j]) - mean[j];
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < kPerThread; ++u) {
            if (pi[u] < 0) continue;
            double s = acc[u];
            for (int t = 0; t < len; ++t) s = fma(chunk[t][pi[u]], chunk[t][pj[u]], s);
            acc[u] = s;
        }
        __syncthreads();
    }
    const double scale = prm.obs_per_year / (double)(nd - 1);
    for (int x = tid; x < kNP * (kNP + 1); x += kCalBlock) {
        (&A[0][0])[x] = 0.0;
        (&V[0][0])[x] = 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < kPerThread; ++u) {
        if (pi[u] < 0) continue;
        const double c = acc[u] * scale;
        A[pi[u]][pj[u]] = c;
        A[pj[u]][pi[u]] = c;
    }
    if (tid < kNP) V[tid][tid] = 1.0;
    __syncthreads();

    // ---- 2. cyclic Jacobi, tournament ordering
    for (int sweep = 0; sweep < kMaxSweeps; ++sweep) {
        // convergence: off-diagonal mass against the total
        double off = 0.0, tot = 0.0;
        for (int x = tid; x < kN * kN; x += kCalBlock) {
            const int i = x / kN, j = x % kN;
            const double a2 = A[i][j] * A[i][j];
            tot += a2;
            if (i != j) off += a2;
        }
        for (int o = kWave / 2; o > 0; o >>= 1) {
            off += __shfl_xor(off, o);
            tot += __shfl_xor(tot, o);
        }
        if ((tid & (kWave - 1)) == 0) {
            red[tid / kWave] = off;
            red2[tid / kWave] = tot;
        }
        __syncthreads();
        double offs = 0.0, tots = 0.0;
        for (int q = 0; q < kCalBlock / kWave; ++q) {
            offs += red[q];
            tots += red2[q];
        }
        __syncthreads();
        if (!(offs > 1e-28 * tots)) break;   // uniform across the workgroup
        for (int r = 0; r < kNP - 1; ++r) {
            if (tid < kNP / 2) {
                int p, q;
                tournament_pair(r, tid, p, q);
                double c = 1.0, s = 0.0;
                const double apq = A[p][q];
                if (q < kN && apq != 0.0) {
                    const double theta = (A[q][q] - A[p][p]) / (2.0 * apq);
                    const double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(1.0 + theta * theta));
                    c = 1.0 / sqrt(1.0 + t * t);
                    s = t * c;
                }
                cs[tid] = c;
                sn[tid] = s;
            }
            __syncthreads();
            // rows p, q of A <- J^T A
            for (int x = tid; x < (kNP / 2) * kNP; x += kCalBlock) {
                const int pr = x / kNP, k = x % kNP;
                int p, q;
                tournament_pair(r, pr, p, q);
                const double c = cs[pr], s = sn[pr];
                const double ap = A[p][k], aq = A[q][k];
                A[p][k] = c * ap - s * aq;
                A[q][k] = s * ap + c * aq;
            }
            __syncthreads();
            // columns p, q of A and V <- A J, V J
            for (int x = tid; x < (kNP / 2) * kNP; x += kCalBlock) {
                const int pr = x / kNP, k = x % kNP;
                int p, q;
                tournament_pair(r, pr, p, q);
                const double c = cs[pr], s = sn[pr];
                const double ap = A[k][p], aq = A[k][q];
                A[k][p] = c * ap - s * aq;
                A[k][q] = s * ap + c * aq;
                const double vp = V[k][p], vq = V[k][q];
                V[k][p] = c * vp - s * vq;
                V[k][q] = s * vp + c * vq;
            }
            __syncthreads();
        }
    }

    // ---- 3. order, signs, outputs
    if (tid == 0) {
        for (int i = 0; i < kN; ++i) order[i] = i;
        for (int i = 1; i < kN; ++i) {  // insertion sort, descending, stable
            const int o = order[i];
            int j = i - 1;
            while (j >= 0 && A[order[j]][order[j]] < A[o][o]) {
                order[j + 1] = order[j];
                --j;
            }
            order[j + 1] = o;
        }
    }
    __syncthreads();
    if (tid < kN) evals_out[(int64_t)w * kN + tid] = A[order[tid]][order[tid]];
    const int F = prm.nfactors;
    if (tid < F) {
        const int col = order[tid];
        double sum = 0.0;
        for (int j = 0; j < kN; ++j) sum += V[j][col];
        double sign = 1.0;
        if (fabs(sum) > 1e-9) {
            sign = sum > 0.0 ? 1.0 : -1.0;
        } else {
            for (int j = 0; j < kN; ++j)
                if (fabs(V[j][col]) > 1e-9) {
                    sign = V[j][col] > 0.0 ? 1.0 : -1.0;
                    break;
                }
        }
        const double lam = A[col][col];
        const double sl = lam > 0.0 ? sqrt(lam) : 0.0;
        double* raw = raw_out + ((int64_t)w * F + tid) * kN;
        for (int j = 0; j < kN; ++j) raw[j] = sign * sl * V[j][col];
        // least-squares polynomial in x = tau / tau_max (normal equations, Gauss-Jordan
        // with partial pivoting; degree <= 5 keeps them well conditioned on [0, 1])
        double* out = loadings_out + ((int64_t)w * F + tid) * kN;
        const int d = prm.degree[tid];
        if (d < 0) {
            for (int j = 0; j < kN; ++j) out[j] = raw[j];
        } else {
            const int m = d + 1;
            double G[kMaxDegree + 1][kMaxDegree + 2];
            for (int a = 0; a < m; ++a)
                for (int b = 0; b <= m; ++b) G[a][b] = 0.0;
            for (int j = 0; j < kN; ++j) {
                const double x = (double)j / (double)(kN - 1);
                double pa = 1.0;
                for (int a = 0; a < m; ++a) {
                    double pb = 1.0;
                    for (int b = 0; b < m; ++b) {
                        G[a][b] += pa * pb;
                        pb *= x;
                    }
                    G[a][m] += pa * raw[j];
                    pa *= x;
                }
            }
            for (int c = 0; c < m; ++c) {
                int piv = c;
                for (int r2 = c + 1; r2 < m; ++r2)
                    if (fabs(G[r2][c]) > fabs(G[piv][c])) piv = r2;
                if (piv != c)
                    for (int b = 0; b <= m; ++b) {
                        const double t = G[c][b];
                        G[c][b] = G[piv][b];
                        G[piv][b] = t;
                    }
                const double inv = 1.0 / G[c][c];
                for (int b = c; b <= m; ++b) G[c][b] *= inv;
                for (int r2 = 0; r2 < m; ++r2) {
                    if (r2 == c) continue;
                    const double f = G[r2][c];
                    for (int b = c; b <= m; ++b) G[r2][b] -= f * G[c][b];
                }
            }
            for (int j = 0; j < kN; ++j) {
                const double x = (double)j / (double)(kN - 1);
                double v = 0.0;
                for (int a = m - 1; a >= 0; --a) v = v * x + G[a][m];
                out[j] = v;
            }
        }
    }
}

}  // namespace

extern "C" {

int hjm_calibrate_pca(const hjm_calibration* c, hjm_calibration_results* r)
{
    if (!c || !r || !c->curves || !r->eigenvalues || !r->volatilities || !r->drifts)
        return HJM_ERR_NULL;
    if (c->window < 3 || c->nwindows < 1 || c->step < 1 || c->nfactors < 1 ||
        c->nfactors > kMaxFactors || !(c->obs_per_year > 0.0) || !(c->dtau > 0.0))
        return HJM_ERR_ARG;
    if (c->nfactors != kFactors) return HJM_ERR_ARG;  // the path kernels take 3 loadings
    if ((int64_t)(c->nwindows - 1) * c->step + c->window > c->ncurves) return HJM_ERR_ARG;
    CalibParams prm{};
    prm.ncurves = c->ncurves;
    prm.window = c->window;
    prm.step = c->step;
    prm.nwindows = c->nwindows;
    prm.nfactors = c->nfactors;
    prm.obs_per_year = c->obs_per_year;
    prm.dtau = c->dtau;
    for (int k = 0; k < kMaxFactors; ++k) prm.degree[k] = -1;
    if (c->fit_degrees)
        for (int k = 0; k < c->nfactors; ++k) {
            if (c->fit_degrees[k] > kMaxDegree) return HJM_ERR_ARG;
            prm.degree[k] = c->fit_degrees[k];
        }
    for (int64_t i = 0; i < c->ncurves * kN; ++i)
        if (!isfinite(c->curves[i])) return HJM_ERR_ARG;

    DeviceGuard guard;
    HJM_TRY(guard.set(c->device));
    const size_t W = (size_t)c->nwindows, F = (size_t)c->nfactors;
    double *d_curves = nullptr, *d_ev = nullptr, *d_ld = nullptr, *d_raw = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = dev_alloc(&d_curves, (size_t)c->ncurves * kN);
    if (rc == HJM_OK) rc = dev_alloc(&d_ev, W * kN);
    if (rc == HJM_OK) rc = dev_alloc(&d_ld, W * F * kN);
    if (rc == HJM_OK) rc = dev_alloc(&d_raw, W * F * kN);
    if (rc == HJM_OK && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
        rc = HJM_ERR_DEVICE;
    if (rc == HJM_OK &&
        hipMemcpy(d_curves, c->curves, (size_t)c->ncurves * kN * sizeof(double),
                  hipMemcpyHostToDevice) != hipSuccess)
        rc = HJM_ERR_DEVICE;
    if (rc == HJM_OK) {
        (void)hipEventRecord(e0, nullptr);
        calib_pca_kernel<<<dim3((unsigned)W), kCalBlock, 0, nullptr>>>(prm, d_curves, d_ev, d_ld, d_raw);
        if (hipGetLastError() != hipSuccess) rc = HJM_ERR_LAUNCH;
        (void)hipEventRecord(e1, nullptr);
    }
    std::vector<double> ld(W * F * kN);
    if (rc == HJM_OK && hipDeviceSynchronize() != hipSuccess) rc = HJM_ERR_LAUNCH;
    if (rc == HJM_OK &&
        (hipMemcpy(r->eigenvalues, d_ev, W * kN * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess ||
         hipMemcpy(ld.data(), d_ld, ld.size() * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess ||
         (r->raw_loadings &&
          hipMemcpy(r->raw_loadings, d_raw, ld.size() * sizeof(double), hipMemcpyDeviceToHost) !=
              hipSuccess)))
        rc = HJM_ERR_DEVICE;
    if (rc == HJM_OK && r->timings_ms) {
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        r->timings_ms[0] = ms;
    }
    if (rc == HJM_OK) {
        memcpy(r->volatilities, ld.data(), ld.size() * sizeof(double));
        for (size_t w = 0; w < W && rc == HJM_OK; ++w)
            rc = hjm_hjm_drift(&ld[w * F * kN], c->dtau, r->drifts + w * kN);
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    dev_free(d_curves);
    dev_free(d_ev);
    dev_free(d_ld);
    dev_free(d_raw);
    return rc;
}

}  // extern "C"
