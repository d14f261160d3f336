/*
 * hjm_oracle.c -- CPU oracle (test infrastructure only; see hjm_oracle.h).
 *
 * Built with -O2 -ffp-contract=off so that every a*b+c below is a rounded multiply
 * followed by a rounded add, as the reference's sequential C++ computes it, and so
 * that the generator's uniforms (which use an explicit fmaf) are bit-reproducible.
 */
#include "hjm_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ===================================================================================
 * Normal generator specification.
 *
 * Class streams (Philox4x32-10: Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
 * easy as 1, 2, 3", SC'11). A run's steps form date blocks of S = dtau/dt steps. Step n
 * of path p lies in block b = n / S at offset s = n % S; its class is g = s % 4 and it is
 * the q = s / 4-th step of class g in the block. Its factor k is shock j = 3q + k of the
 * stream (p, b, g). Call c of that stream is one Philox block with counter (b, p, g, c)
 * and key (seed mod 2^32, seed >> 32), giving shocks 4c .. 4c + 3:
 *   x0..x3 = philox(ctr, key)
 *   (4c, 4c + 1)     = r cos, r sin of (x0, x1)
 *   (4c + 2, 4c + 3) = r cos, r sin of (x2, x3)
 * with, for a pair (a, b): u = fma(float(a), 2^-32, 2^-33) in (0, 1] (float(a) is a rounded
 * to 24 bits, round to nearest even), t = float(b) 2^-32 in [0, 1] turns, r = sqrt(-2 ln u),
 * shocks r cos(2 pi t) and r sin(2 pi t). u and t are exact FP32 values, reproduced here
 * bit for bit; the transform itself is evaluated in FP64 and rounded once to FP32. The
 * product runs the transform on the GPU's log2 / sqrt / sin / cos instructions, whose
 * results differ from these by at most 2.6e-7 |z| + 8.8e-7 (the bound the GPU tests
 * check). The reference draws cuRAND XORWOW normals instead (simulation_cpu.h:83-90,
 * simulation_gpu.cu:1011-1013); that sequence is closed library behaviour and is replaced,
 * not reproduced.
 * =================================================================================== */

#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += PHILOX_W0;
            k1 += PHILOX_W1;
        }
        uint64_t p0 = (uint64_t)PHILOX_M0 * c0;
        uint64_t p1 = (uint64_t)PHILOX_M1 * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

/* (z0, z1) = r cos, r sin of the Philox pair (a, b): u = fma(float(a), 2^-32, 2^-33),
 * t = float(b) 2^-32, r = sqrt(-2 ln u), angle 2 pi t; FP64 transform rounded to FP32. */
void orc_box_muller(uint32_t a, uint32_t b, float* z0, float* z1)
{
    const float u = fmaf((float)a, 0x1p-32f, 0x1p-33f);
    const float t = (float)b * 0x1p-32f;
    const double r = sqrt(-2.0 * log((double)u));
    const double ang = 6.283185307179586476925 * (double)t;
    *z0 = (float)(r * cos(ang));
    *z1 = (float)(r * sin(ang));
}

/* Shocks 4 call .. 4 call + 3 of the class stream (path, block, cls): Philox4x32-10 with
 * counter (block, path, cls, call) and key = seed; Box-Muller on (x0, x1) and (x2, x3). */
void orc_stream_call(uint64_t seed, uint32_t path, uint32_t block, uint32_t cls, uint32_t call,
                     float z[4])
{
    uint32_t ctr[4] = {block, path, cls, call};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t x[4];
    orc_philox4x32_10(ctr, key, x);
    orc_box_muller(x[0], x[1], &z[0], &z[1]);
    orc_box_muller(x[2], x[3], &z[2], &z[3]);
}

/* Step n of a run whose date blocks are `stride` steps long: block b = n / stride, offset
 * s = n % stride, class g = s % 4, q = s / 4; its factor k is shock j = 3 q + k of the
 * stream (path, b, g), i.e. element j % 4 of call j / 4. */
void orc_step_normals(uint64_t seed, uint32_t path, uint32_t step, uint32_t stride, float z[3])
{
    uint32_t b = step / stride, s = step % stride, g = s % 4u, q = s / 4u;
    for (int k = 0; k < 3; ++k) {
        uint32_t j = 3u * q + (uint32_t)k;
        float c[4];
        orc_stream_call(seed, path, b, g, j / 4u, c);
        z[k] = c[j % 4u];
    }
}

/* Layout [path][step][3], as the reference stores its normals (simulation_cpu.h:205-207).
 * Walks each stream (path, block, class) once: every call is made once and its four
 * shocks are scattered to their steps (the same values as orc_step_normals). */
void orc_fill_normals(uint64_t seed, int64_t path0, int npaths, int nsteps, int stride,
                      double* out)
{
    const int nblocks = (nsteps + stride - 1) / stride;
    for (int p = 0; p < npaths; ++p) {
        for (int b = 0; b < nblocks; ++b) {
            for (int g = 0; g < 4 && g < stride; ++g) {
                const int nq = (stride - g + 3) / 4;  /* steps of class g in a block */
                const int ncall = (3 * nq + 3) / 4;
                for (int c = 0; c < ncall; ++c) {
                    float z[4];
                    orc_stream_call(seed, (uint32_t)(path0 + p), (uint32_t)b, (uint32_t)g,
                                    (uint32_t)c, z);
                    for (int e = 0; e < 4; ++e) {
                        const int j = 4 * c + e, q = j / 3, k = j % 3;
                        if (q >= nq) break;
                        const int n = b * stride + 4 * q + g;
                        if (n >= nsteps) continue;
                        out[((int64_t)p * nsteps + n) * 3 + k] = z[e];
                    }
                }
            }
        }
    }
}

/* ===================================================================================
 * Market data. The reference takes a drift table as input (simulation_gpu_testing.cpp
 * :66-71); it equals the HJM no-arbitrage drift m(tau) = sum_k sigma_k(tau) *
 * int_0^tau sigma_k(u) du of its volatility table (:47-63) with the integral taken by
 * the cumulative Simpson rule below (checked to the table's 7 printed decimals).
 * =================================================================================== */

/* int_0^{x_j} y for j = 0..n-1 on a uniform grid of step h:
 * even j: composite Simpson 1/3; odd j >= 3: Simpson 1/3 to j-3, Simpson 3/8 on the
 * last three intervals; j = 1: the quadratic through points 0, 1, 2 over [x0, x1]. */
void orc_cumulative_simpson(const double* y, int n, double h, double* out)
{
    out[0] = 0.0;
    for (int j = 1; j < n; ++j) {
        double s = 0.0;
        if (j == 1) {
            s = (n > 2) ? h / 12.0 * (5.0 * y[0] + 8.0 * y[1] - y[2]) : 0.5 * h * (y[0] + y[1]);
        } else {
            int even_end = (j % 2 == 0) ? j : j - 3;
            for (int i = 0; i < even_end; i += 2) s += h / 3.0 * (y[i] + 4.0 * y[i + 1] + y[i + 2]);
            if (j % 2 == 1)
                s += 3.0 * h / 8.0 * (y[j - 3] + 3.0 * y[j - 2] + 3.0 * y[j - 1] + y[j]);
        }
        out[j] = s;
    }
}

void orc_hjm_drift(const double* vols, int nfactors, int ntenors, double h, double* drift)
{
    double* integ = (double*)malloc(sizeof(double) * ntenors);
    for (int j = 0; j < ntenors; ++j) drift[j] = 0.0;
    for (int k = 0; k < nfactors; ++k) {
        orc_cumulative_simpson(vols + k * ntenors, ntenors, h, integ);
        for (int j = 0; j < ntenors; ++j) drift[j] += vols[k * ntenors + j] * integ[j];
    }
    free(integ);
}

/* Nelson-Siegel instantaneous forward curve f(tau) = b0 + b1 e^{-tau/l} + b2 (tau/l) e^{-tau/l}. */
void orc_nelson_siegel(double b0, double b1, double b2, double lam, int ntenors, double h,
                       double* curve)
{
    for (int j = 0; j < ntenors; ++j) {
        double x = j * h / lam;
        double e = exp(-x);
        curve[j] = b0 + b1 * e + b2 * x * e;
    }
}

/* The three parametric volatility factors of the benchmark configurations (factor-major). */
void orc_poly_vols(int ntenors, double h, double* vols)
{
    for (int j = 0; j < ntenors; ++j) {
        double t = j * h;
        vols[j] = 0.0064;
        vols[ntenors + j] = -0.0036 + 0.0006 * t - 2e-5 * t * t;
        vols[2 * ntenors + j] = -0.0047 + 0.0018 * t - 1.5e-4 * t * t + 3e-6 * t * t * t;
    }
}

/* ===================================================================================
 * Reference semantics: simulation_cpu.h restated.
 * =================================================================================== */

/* simulation_cpu.h:37-54, same operation order. */
static double musiela_sde(double drift, double vol0, double vol1, double vol2, double phi0,
                          double phi1, double phi2, double dF, double rate0, double dtau,
                          double dt)
{
    double vol_sum = vol0 * phi0;
    vol_sum += vol1 * phi1;
    vol_sum += vol2 * phi2;
    vol_sum *= sqrt(dt);
    double dfbar = drift * dt;
    dfbar += vol_sum;
    dfbar += (dF / dtau) * dt;
    return rate0 + dfbar;
}

/* simulation_cpu.h:96-108: forward difference, backward at the last tenor. */
static double dfau(int t, const double* rates)
{
    return (t == ORC_TENORS - 1) ? rates[t] - rates[t - 1] : rates[t + 1] - rates[t];
}

/* Time grid: the reference truncates stride = (int)(dtau/dt) and steps = (int)(expiry/dt)
 * (simulation_cpu.h:200, :384, :409); both ratios are rounded here and must be integral
 * to 1e-6 relative, the path must be a whole number of blocks, and the last numeraire
 * slot must lie on the grid (checks the reference lacks). Identical to the reference
 * whenever dtau/dt and expiry/dt are integers, as in every configuration. */
static int ref_grid(double expiry, double dtau, double dt, int* stride, int* nsteps, int* ndates)
{
    if (!(dt > 0.0) || !(dtau > 0.0) || !(expiry > 0.0)) return 2;
    double rs = dtau / dt, rS = expiry / dt;
    if (rs > 1e7 || rS > 1e9) return 3;
    int st = (int)llround(rs);
    int S = (int)llround(rS);
    if (st <= 0 || fabs(rs - st) > 1e-6 * st || fabs(rS - S) > 1e-6 * S) return 3;
    if (S < st || S % st != 0) return 3;
    int D = S / st;
    if (D > ORC_TENORS - 1) return 3;
    *stride = st;
    *nsteps = S;
    *ndates = D;
    return 0;
}

/* One path of simulation_cpu.h:188-262 driven by the loop at :413-436. Slots that the
 * reference never writes (0 and beyond the last date) are left at 0: the reference's
 * malloc'd numeraire buffer is fresh, zero-filled memory at every size it was run at. */
static void ref_path_one(const double* spot, const double* drift, const double* vols,
                         double dtau, double dt, int stride, int nsteps, int ndates,
                         const double* nrm, double* nx, double* ny)
{
    double cur[ORC_TENORS], nxt[ORC_TENORS], acc[ORC_TENORS];
    for (int t = 0; t < ORC_TENORS; ++t) {
        cur[t] = spot[t];
        acc[t] = 0.0;
        nx[t] = 0.0;
        ny[t] = 0.0;
    }
    for (int path = 0; path < nsteps; path += stride) {
        for (int s = 0; s < stride; ++s) {
            const double* phi = nrm + (int64_t)(path + s) * 3;
            for (int t = 0; t < ORC_TENORS; ++t) {
                double rate;
                if (path == 0) {
                    rate = spot[t];
                } else {
                    rate = musiela_sde(drift[t], vols[t], vols[ORC_TENORS + t],
                                       vols[2 * ORC_TENORS + t], phi[0], phi[1], phi[2],
                                       dfau(t, cur), cur[t], dtau, dt);
                }
                acc[t] += rate;
                nxt[t] = rate;
            }
            memcpy(cur, nxt, sizeof(cur));
        }
        int l = (path + stride) / stride;
        nx[l] = nxt[l];
        ny[l] = exp(-acc[l] * dt);
    }
    (void)ndates;
}

int orc_reference_paths(const double* spot, const double* drift, const double* vols,
                        double expiry, double dtau, double dt, int npaths,
                        const double* normals, double* num_x, double* num_y)
{
    int stride, S, D;
    int rc = ref_grid(expiry, dtau, dt, &stride, &S, &D);
    if (rc) return rc;
    if (npaths <= 0) return 2;
    for (int p = 0; p < npaths; ++p)
        ref_path_one(spot, drift, vols, dtau, dt, stride, S, D, normals + (int64_t)p * S * 3,
                     num_x + (int64_t)p * ORC_TENORS, num_y + (int64_t)p * ORC_TENORS);
    return 0;
}

/* simulation_cpu.h:269-296 (payer swap; t = 51 out-of-bounds write omitted). */
void orc_reference_exposure(const double* num_x, const double* num_y, const double* accrual,
                            double notional, double K, double dtau, int npaths,
                            double* exposures)
{
    double m = 1.0 / dtau;
    for (int p = 0; p < npaths; ++p) {
        const double* x = num_x + (int64_t)p * ORC_TENORS;
        const double* y = num_y + (int64_t)p * ORC_TENORS;
        double cf[ORC_TENORS];
        for (int t = 0; t < ORC_TENORS; ++t) {
            double libor = m * (exp(x[t] / m) - 1.0);
            cf[t] = y[t] * notional * accrual[t] * (libor - K);
        }
        double* e = exposures + (int64_t)p * ORC_TENORS;
        for (int t = 0; t < ORC_TENORS; ++t) {
            double sum = 0.0;
            for (int i = t + 1; i < ORC_TENORS; ++i) sum += cf[i];
            e[t] = (sum > 0.0) ? sum : 0.0;
        }
    }
}

/* simulation_cpu.h:301-343: EE = (1/N) * column sums (cblas_dgemv with alpha = 1/N). */
void orc_expected_exposure(const double* exposures, int npaths, int ndates, double* ee)
{
    double alpha = 1.0 / (double)npaths;
    for (int t = 0; t < ndates; ++t) {
        double s = 0.0;
        for (int p = 0; p < npaths; ++p) s += exposures[(int64_t)p * ndates + t];
        ee[t] = alpha * s;
    }
}

/* calculateExposureCPU<double> (simulation_cpu.h:349-458) with injected normals. */
int orc_calculate_exposure_cpu(double* expected_exposure, const double* spot,
                               const double* drift, const double* vols,
                               const double* accrual, double notional, double K,
                               double expiry, double dtau, double dt, int npaths,
                               const double* normals, double* exposures)
{
    size_t n = (size_t)npaths * ORC_TENORS;
    double* nx = (double*)malloc(n * sizeof(double));
    double* ny = (double*)malloc(n * sizeof(double));
    double* ex = exposures ? exposures : (double*)malloc(n * sizeof(double));
    int rc = orc_reference_paths(spot, drift, vols, expiry, dtau, dt, npaths, normals, nx, ny);
    if (!rc) {
        orc_reference_exposure(nx, ny, accrual, notional, K, dtau, npaths, ex);
        orc_expected_exposure(ex, npaths, ORC_TENORS, expected_exposure);
    }
    free(nx);
    free(ny);
    if (!exposures) free(ex);
    return rc;
}

/* ===================================================================================
 * Statistics (no reference counterpart; doc/TODO:14 names CVA as the goal).
 * =================================================================================== */

/* CVA = (1-R) sum_i 0.5 (EE(t_{i-1}) + EE(t_i)) (S(t_{i-1}) - S(t_i)), S(t) = e^{-h t},
 * t_i = i * date_dt: trapezoid in exposure, exact default probability per interval.
 * Returned as weights w_i so that CVA = sum_i w_i EE_i (and per path likewise). */
void orc_cva_weights(int ndates, double date_dt, double hazard, double recovery, double* w)
{
    for (int i = 0; i < ndates; ++i) w[i] = 0.0;
    for (int i = 1; i < ndates; ++i) {
        double pd = exp(-hazard * (i - 1) * date_dt) - exp(-hazard * i * date_dt);
        w[i - 1] += 0.5 * (1.0 - recovery) * pd;
        w[i] += 0.5 * (1.0 - recovery) * pd;
    }
}

/* Inverted-CDF quantile: the k-th smallest sample, k = ceil(q n) clamped to [1, n]. */
int64_t orc_quantile_rank(int64_t n, double q)
{
    double v = q * (double)n;
    int64_t k = (int64_t)ceil(v);
    if (k < 1) k = 1;
    if (k > n) k = n;
    return k;
}

static int cmp_double(const void* a, const void* b)
{
    double x = *(const double*)a, y = *(const double*)b;
    return (x > y) - (x < y);
}

void orc_stats(const double* exposures, int npaths, int ndates, const double* w, double q,
               double* ee, double* pfe, double* cva, double* cva_se)
{
    orc_expected_exposure(exposures, npaths, ndates, ee);
    double* col = (double*)malloc(sizeof(double) * npaths);
    int64_t k = orc_quantile_rank(npaths, q);
    for (int t = 0; t < ndates; ++t) {
        for (int p = 0; p < npaths; ++p) col[p] = exposures[(int64_t)p * ndates + t];
        qsort(col, npaths, sizeof(double), cmp_double);
        pfe[t] = col[k - 1];
    }
    free(col);
    double c = 0.0;
    for (int t = 0; t < ndates; ++t) c += w[t] * ee[t];
    double s1 = 0.0, s2 = 0.0;
    for (int p = 0; p < npaths; ++p) {
        double v = 0.0;
        for (int t = 0; t < ndates; ++t) v += w[t] * exposures[(int64_t)p * ndates + t];
        s1 += v;
        s2 += v * v;
    }
    *cva = c;
    double var = (npaths > 1) ? (s2 - s1 * s1 / npaths) / (npaths - 1) : 0.0;
    *cva_se = sqrt((var > 0.0 ? var : 0.0) / npaths);
}

/* ===================================================================================
 * Pathwise adjoint of the per-path CVA sample J = sum_t w_t max(V_t, 0).
 *
 * Forward (per path): f(0) = spot, A(0) = stride * spot (frozen block), for evolution
 * steps n = 1..N (N = (D-1) stride, normals of global step stride + n - 1):
 *   f(n) = T f(n-1) + m dt + sqrt(dt) V phi(n),  A(n) = A(n-1) + f(n),
 * with T = I + (dt/dtau) Dfwd. Slot i = 1..D records x_i = f_i(n_i), A_i(n_i),
 * n_i = (i-1) stride; CF_i = e^{-A_i dt} N acc_i (m (e^{x_i/m} - 1) - K).
 * Seeds: u_i = sum_{t<i} w_t 1[V_t > 0], gx_i = u_i e^{-A_i dt} N acc_i e^{x_i/m},
 * gA_i = -u_i dt CF_i. Backward: mu(n) = mu(n+1) + sum_{n_i = n} gA_i e_i,
 * lambda(n) = T^T lambda(n+1) + mu(n) + sum_{n_i = n} gx_i e_i;
 * dJ/dV_kj = sum_n lambda_j(n) sqrt(dt) phi_k(n), dJ/dm_j = dt sum_n lambda_j(n),
 * dJ/dspot = T^T lambda(1) + gx_1 e_1 + stride mu(0), mu(0) = mu(1) + gA_1 e_1.
 * The step Jacobian T does not depend on the state, so no forward states are needed
 * between dates.
 * =================================================================================== */
static void apply_TT(double* lam, double c)
{
    /* (T^T l)_j = (1-c) l_j + c l_{j-1} (1 <= j <= 48), j=0: (1-c) l_0,
     * j=49: (1-c) l_49 + c l_48 - c l_50, j=50: (1+c) l_50 + c l_49. */
    double out[ORC_TENORS];
    out[0] = (1.0 - c) * lam[0];
    for (int j = 1; j <= 48; ++j) out[j] = (1.0 - c) * lam[j] + c * lam[j - 1];
    out[49] = (1.0 - c) * lam[49] + c * lam[48] - c * lam[50];
    out[50] = (1.0 + c) * lam[50] + c * lam[49];
    memcpy(lam, out, sizeof(out));
}

int orc_reference_cva_greeks(const double* spot, const double* drift, const double* vols,
                             const double* accrual, double notional, double K, double expiry,
                             double dtau, double dt, int npaths, const double* normals,
                             const double* w, double* cva, double* d_spot, double* d_vol,
                             double* d_drift)
{
    int stride, S, D;
    int rc = ref_grid(expiry, dtau, dt, &stride, &S, &D);
    if (rc) return rc;
    if (npaths <= 0) return 2;
    const int N = (D - 1) * stride;
    const double c = dt / dtau, sq = sqrt(dt), m = 1.0 / dtau;
    double gs[ORC_TENORS] = {0}, gv[3 * ORC_TENORS] = {0}, gm[ORC_TENORS] = {0}, J = 0.0;
    for (int p = 0; p < npaths; ++p) {
        const double* nrm = normals + (int64_t)p * S * 3;
        double nx[ORC_TENORS], ny[ORC_TENORS];
        ref_path_one(spot, drift, vols, dtau, dt, stride, S, D, nrm, nx, ny);
        double cf[ORC_TENORS] = {0}, V[ORC_TENORS] = {0};
        for (int i = 1; i <= D; ++i) {
            double libor = m * (exp(nx[i] / m) - 1.0);
            cf[i] = ny[i] * notional * accrual[i] * (libor - K);
        }
        double acc = 0.0;
        for (int t = ORC_TENORS - 1; t >= 0; --t) {
            if (t + 1 < ORC_TENORS) acc += cf[t + 1];
            V[t] = acc;
            if (acc > 0.0) J += w[t] * acc;
        }
        double gx[ORC_TENORS] = {0}, gA[ORC_TENORS] = {0};
        double u = 0.0;
        for (int i = 1; i <= D; ++i) {
            if (V[i - 1] > 0.0) u += w[i - 1];
            gx[i] = u * ny[i] * notional * accrual[i] * exp(nx[i] / m);
            gA[i] = -u * dt * cf[i];
        }
        double lam[ORC_TENORS] = {0}, mu[ORC_TENORS] = {0};
        for (int n = N; n >= 1; --n) {
            if (n % stride == 0) {
                int i = n / stride + 1;
                lam[i] += gx[i];
                mu[i] += gA[i];
            }
            const double* phi = nrm + (int64_t)(stride + n - 1) * 3;
            for (int j = 0; j < ORC_TENORS; ++j) {
                lam[j] += mu[j];
                gv[j] += lam[j] * sq * phi[0];
                gv[ORC_TENORS + j] += lam[j] * sq * phi[1];
                gv[2 * ORC_TENORS + j] += lam[j] * sq * phi[2];
                gm[j] += dt * lam[j];
            }
            apply_TT(lam, c);
        }
        lam[1] += gx[1];
        mu[1] += gA[1];
        for (int j = 0; j < ORC_TENORS; ++j) gs[j] += lam[j] + stride * mu[j];
    }
    const double inv = 1.0 / npaths;
    if (cva) *cva = J * inv;
    for (int j = 0; j < ORC_TENORS; ++j) {
        if (d_spot) d_spot[j] = gs[j] * inv;
        if (d_drift) d_drift[j] = gm[j] * inv;
    }
    if (d_vol)
        for (int j = 0; j < 3 * ORC_TENORS; ++j) d_vol[j] = gv[j] * inv;
    return 0;
}

void orc_drift_vjp(const double* vols, double h, const double* g, double* out)
{
    /* m_j = sum_k s_kj I_kj, I_k = W s_k (W: cumulative Simpson weights, linear).
     * dm_j/ds_kl = delta_jl I_kj + s_kj W_jl. */
    double e[ORC_TENORS], col[ORC_TENORS];
    for (int k = 0; k < 3; ++k) {
        const double* s = vols + k * ORC_TENORS;
        double I[ORC_TENORS];
        orc_cumulative_simpson(s, ORC_TENORS, h, I);
        for (int l = 0; l < ORC_TENORS; ++l) {
            memset(e, 0, sizeof(e));
            e[l] = 1.0;
            orc_cumulative_simpson(e, ORC_TENORS, h, col); /* col[j] = W_jl */
            double acc = g[l] * I[l];
            for (int j = 0; j < ORC_TENORS; ++j) acc += g[j] * s[j] * col[j];
            out[k * ORC_TENORS + l] += acc;
        }
    }
}

/* ===================================================================================
 * Portfolio semantics (declared in hjm_oracle.h).
 * =================================================================================== */
static void bond_prices(const double* f, double dtau, int kmax, double* P)
{
    double I = 0.0;
    for (int k = 1; k <= kmax; ++k) {
        if (k <= ORC_TENORS - 1) I += 0.5 * dtau * (f[k - 1] + f[k]);
        else I += dtau * f[ORC_TENORS - 1];
        P[k] = exp(-I);
    }
    P[0] = 1.0;
}

int orc_portfolio_exposure(const double* spot, const double* drift, const double* vols,
                           double horizon, double dtau, double dt, int npaths,
                           const double* normals, int nswaps, const int* periods,
                           const double* sign_notional, const double* fixed_rate,
                           const int* set_of, int nsets, double* exposures, double* curves,
                           double* log_bank)
{
    if (!(dt > 0.0) || !(dtau > 0.0) || !(horizon > 0.0) || npaths <= 0 || nsets <= 0) return 2;
    double rs = dtau / dt, rS = horizon / dt;
    int st = (int)llround(rs), S = (int)llround(rS);
    if (st <= 0 || fabs(rs - st) > 1e-6 * st || fabs(rS - S) > 1e-6 * S || S % st) return 3;
    const int E = S / st;  /* last date index; dates 0..E */
    for (int s = 0; s < nswaps; ++s)
        if (periods[s] < 1 || periods[s] > E || set_of[s] < 0 || set_of[s] >= nsets) return 2;
    const int nd = E + 1;
    double* P = (double*)malloc(sizeof(double) * (E + 2));
    double* V = (double*)malloc(sizeof(double) * nsets);
    for (int p = 0; p < npaths; ++p) {
        const double* nrm = normals + (int64_t)p * S * 3;
        double cur[ORC_TENORS], nxt[ORC_TENORS];
        memcpy(cur, spot, sizeof(cur));
        double lnB = 0.0;
        for (int n = 0; n <= S; ++n) {
            if (n % st == 0) {
                const int e = n / st;
                if (curves) memcpy(curves + ((int64_t)p * nd + e) * ORC_TENORS, cur, sizeof(cur));
                if (log_bank) log_bank[(int64_t)p * nd + e] = lnB;
                bond_prices(cur, dtau, E - e, P);
                for (int g = 0; g < nsets; ++g) V[g] = 0.0;
                for (int s = 0; s < nswaps; ++s) {
                    const int k = periods[s] - e;
                    if (k <= 0) continue;
                    double ann = 0.0;
                    for (int j = 1; j <= k; ++j) ann += P[j];
                    V[set_of[s]] += sign_notional[s] * (1.0 - P[k] - fixed_rate[s] * dtau * ann);
                }
                const double disc = exp(-lnB);
                for (int g = 0; g < nsets; ++g)
                    exposures[((int64_t)p * nsets + g) * nd + e] = (V[g] > 0.0 ? V[g] : 0.0) * disc;
            }
            if (n == S) break;
            lnB += cur[0] * dt;
            const double* phi = nrm + (int64_t)n * 3;
            for (int t = 0; t < ORC_TENORS; ++t)
                nxt[t] = musiela_sde(drift[t], vols[t], vols[ORC_TENORS + t],
                                     vols[2 * ORC_TENORS + t], phi[0], phi[1], phi[2],
                                     dfau(t, cur), cur[t], dtau, dt);
            memcpy(cur, nxt, sizeof(cur));
        }
    }
    free(P);
    free(V);
    return 0;
}

/* Portfolio adjoint. Per path J = sum_g sum_e w_ge max(V_ge, 0) e^{-lnB_e}.
 * At date e: dJ/dV_ge = w_ge 1[V_ge > 0] D_e (D_e = e^{-lnB_e}), dJ/dlnB_e = -sum_g w_ge E_ge;
 * dJ/dP_k = -sum_g dJ/dV_ge dV_ge/dP_k with V = sum_s sN (1 - P_n - K dtau sum_{j<=n} P_j);
 * dJ/dI_k = -P_k dJ/dP_k; I_k = sum_{l<k} dtau/2 (f_l + f_{l+1}) (k <= 50), flat beyond:
 * dJ/df_l = dtau/2 (S_{l+1} [l <= 49] + S_l [l >= 1]) + [l = 50] dtau sum_{k>50} (k-50) iota_k,
 * S_m = sum_{k >= m} iota_k. lnB_e = dt sum_{n < n_e} f_0(n).
 * Backward: lambda(n) = T^T lambda(n+1) + gbar_e [n = n_e] + mu(n) e_0,
 * mu(n) = dt sum_{e: n_e > n} dJ/dlnB_e; dJ/dV_kj = sum_n lambda_j(n+1) sqrt(dt) phi_k(n),
 * dJ/dm_j = dt sum_n lambda_j(n+1), dJ/dspot = lambda(0). */
int orc_portfolio_cva_greeks(const double* spot, const double* drift, const double* vols,
                             double horizon, double dtau, double dt, int npaths,
                             const double* normals, int nswaps, const int* periods,
                             const double* sign_notional, const double* fixed_rate,
                             const int* set_of, int nsets, const double* weights,
                             double* cva, double* d_spot, double* d_vol, double* d_drift)
{
    if (!(dt > 0.0) || !(dtau > 0.0) || !(horizon > 0.0) || npaths <= 0 || nsets <= 0) return 2;
    double rs = dtau / dt, rS = horizon / dt;
    int st = (int)llround(rs), S = (int)llround(rS);
    if (st <= 0 || fabs(rs - st) > 1e-6 * st || fabs(rS - S) > 1e-6 * S || S % st) return 3;
    const int E = S / st, nd = E + 1;
    const double c = dt / dtau, sq = sqrt(dt);
    double* ex = (double*)malloc(sizeof(double) * nsets * nd);
    double* cv = (double*)malloc(sizeof(double) * nd * ORC_TENORS);
    double* lb = (double*)malloc(sizeof(double) * nd);
    double* gbar = (double*)malloc(sizeof(double) * nd * ORC_TENORS);
    double* beta = (double*)malloc(sizeof(double) * nd);
    double* P = (double*)malloc(sizeof(double) * (E + 2));
    double* dP = (double*)malloc(sizeof(double) * (E + 2));
    double* V = (double*)malloc(sizeof(double) * nsets);
    double gs[ORC_TENORS] = {0}, gv[3 * ORC_TENORS] = {0}, gm[ORC_TENORS] = {0}, J = 0.0;
    for (int p = 0; p < npaths; ++p) {
        const double* nrm = normals + (int64_t)p * S * 3;
        int rc = orc_portfolio_exposure(spot, drift, vols, horizon, dtau, dt, 1, nrm, nswaps,
                                        periods, sign_notional, fixed_rate, set_of, nsets, ex,
                                        cv, lb);
        if (rc) return rc;
        for (int e = 0; e < nd; ++e) {
            const double* f = cv + e * ORC_TENORS;
            const double D = exp(-lb[e]);
            const int kmax = E - e;
            bond_prices(f, dtau, kmax, P);
            for (int g = 0; g < nsets; ++g) V[g] = 0.0;
            for (int s = 0; s < nswaps; ++s) {
                const int k = periods[s] - e;
                if (k <= 0) continue;
                double ann = 0.0;
                for (int j = 1; j <= k; ++j) ann += P[j];
                V[set_of[s]] += sign_notional[s] * (1.0 - P[k] - fixed_rate[s] * dtau * ann);
            }
            double b = 0.0;
            for (int k = 0; k <= kmax; ++k) dP[k] = 0.0;
            for (int g = 0; g < nsets; ++g) {
                const double w = weights[g * nd + e];
                if (V[g] > 0.0) {
                    J += w * V[g] * D;
                    b -= w * V[g] * D;
                }
            }
            for (int s = 0; s < nswaps; ++s) {
                const int k = periods[s] - e;
                if (k <= 0) continue;
                const int g = set_of[s];
                if (!(V[g] > 0.0)) continue;
                const double dv = weights[g * nd + e] * D * sign_notional[s];
                dP[k] -= dv;
                for (int j = 1; j <= k; ++j) dP[j] -= dv * fixed_rate[s] * dtau;
            }
            /* iota_k = dJ/dI_k = -P_k dJ/dP_k ; suffix sums */
            double suffix[64] = {0};
            for (int k = kmax; k >= 1; --k) suffix[k] = suffix[k + 1] + (-P[k] * dP[k]);
            double* gb = gbar + e * ORC_TENORS;
            for (int l = 0; l < ORC_TENORS; ++l) {
                double v = 0.0;
                const int lim = kmax < ORC_TENORS - 1 ? kmax : ORC_TENORS - 1;
                if (l <= 49 && l + 1 <= lim) v += 0.5 * dtau * suffix[l + 1];
                if (l >= 1 && l <= lim) v += 0.5 * dtau * suffix[l];
                if (l == ORC_TENORS - 1)
                    for (int k = ORC_TENORS; k <= kmax; ++k) v += dtau * (k - 50) * (-P[k] * dP[k]);
                /* intervals up to node lim only: I_k for k <= 50 uses f_0..f_k */
                gb[l] = v;
            }
            beta[e] = b;
        }
        /* backward sweep */
        double lam[ORC_TENORS], mu0 = 0.0;
        memcpy(lam, gbar + E * ORC_TENORS, sizeof(lam));
        for (int n = S - 1; n >= 0; --n) {
            const double* phi = nrm + (int64_t)n * 3;
            for (int j = 0; j < ORC_TENORS; ++j) {
                gv[j] += lam[j] * sq * phi[0];
                gv[ORC_TENORS + j] += lam[j] * sq * phi[1];
                gv[2 * ORC_TENORS + j] += lam[j] * sq * phi[2];
                gm[j] += dt * lam[j];
            }
            apply_TT(lam, c);
            if ((n + 1) % st == 0) mu0 += dt * beta[(n + 1) / st];
            lam[0] += mu0;
            if (n % st == 0) {
                const double* gb = gbar + (n / st) * ORC_TENORS;
                for (int j = 0; j < ORC_TENORS; ++j) lam[j] += gb[j];
            }
        }
        for (int j = 0; j < ORC_TENORS; ++j) gs[j] += lam[j];
    }
    const double inv = 1.0 / npaths;
    if (cva) *cva = J * inv;
    for (int j = 0; j < ORC_TENORS; ++j) {
        if (d_spot) d_spot[j] = gs[j] * inv;
        if (d_drift) d_drift[j] = gm[j] * inv;
    }
    if (d_vol)
        for (int j = 0; j < 3 * ORC_TENORS; ++j) d_vol[j] = gv[j] * inv;
    free(ex); free(cv); free(lb); free(gbar); free(beta); free(P); free(dP); free(V);
    return 0;
}

/* ===================================================================================
 * CPU baseline: the path loop of simulation_cpu.h:407-440 timed alone, on nthreads.
 * =================================================================================== */
double orc_time_reference_paths(const double* spot, const double* drift, const double* vols,
                                double expiry, double dtau, double dt, int npaths,
                                const double* normals, double* num_x, double* num_y,
                                int nthreads)
{
    int stride, S, D;
    if (ref_grid(expiry, dtau, dt, &stride, &S, &D) || npaths <= 0) return -1.0;
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
#ifdef _OPENMP
#pragma omp parallel for num_threads(nthreads) schedule(static)
#endif
    for (int p = 0; p < npaths; ++p)
        ref_path_one(spot, drift, vols, dtau, dt, stride, S, D, normals + (int64_t)p * S * 3,
                     num_x + (int64_t)p * ORC_TENORS, num_y + (int64_t)p * ORC_TENORS);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    (void)nthreads;
    return (t1.tv_sec - t0.tv_sec) * 1e3 + (t1.tv_nsec - t0.tv_nsec) * 1e-6;
}
