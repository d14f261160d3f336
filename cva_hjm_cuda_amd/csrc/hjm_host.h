// hjm_host.h -- host-side helpers shared by the C ABI translation units.
#pragma once
#include <math.h>

#include <vector>
#include <stdint.h>

#include "hjm_common.h"

#define HJM_TRY(expr)                            \
    do {                                         \
        const int _rc = (expr);                  \
        if (_rc != HJM_OK) return _rc;           \
    } while (0)

#define HJM_HIP(expr, code)                      \
    do {                                         \
        if ((expr) != hipSuccess) return (code); \
    } while (0)

namespace hjm_host {
using namespace hjm;

// Restores the caller's current device on scope exit.
struct DeviceGuard {
    int prev = -1;
    int set(int dev)
    {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return HJM_ERR_DEVICE;
        if (dev < 0 || dev >= n) return HJM_ERR_ARG;
        if (hipSetDevice(dev) != hipSuccess) return HJM_ERR_DEVICE;
        // a non-sticky error left by an earlier, unrelated call must not be reported by
        // this call's launch checks
        (void)hipGetLastError();
        return HJM_OK;
    }
    ~DeviceGuard()
    {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

template <typename T>
inline int dev_alloc(T** p, size_t count)
{
    *p = nullptr;
    if (count == 0) return HJM_OK;
    if (hipMalloc((void**)p, count * sizeof(T)) != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return HJM_ERR_ALLOC;
    }
    return HJM_OK;
}

template <typename T>
inline void dev_free(T*& p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}


inline int check_market(const hjm_market* m)
{
    if (!m || !m->spot_rates || !m->drifts || !m->volatilities || !m->accrual) return HJM_ERR_NULL;
    for (int j = 0; j < kTenors; ++j) {
        if (!isfinite(m->spot_rates[j]) || !isfinite(m->drifts[j]) || !isfinite(m->accrual[j]))
            return HJM_ERR_ARG;
        for (int k = 0; k < kFactors; ++k)
            if (!isfinite(m->volatilities[k * kTenors + j])) return HJM_ERR_ARG;
    }
    return HJM_OK;
}

template <typename Real>
inline void fill_consts(RefStepConsts<Real>& k, const hjm_market* m, double dt, double dtau)
{
    for (int j = 0; j < kTenors; ++j) {
        k.ddt[j] = (Real)(m->drifts[j] * dt);
        k.v0[j] = (Real)m->volatilities[j];
        k.v1[j] = (Real)m->volatilities[kTenors + j];
        k.v2[j] = (Real)m->volatilities[2 * kTenors + j];
        k.spot[j] = (Real)m->spot_rates[j];
        k.drift[j] = (Real)m->drifts[j];
    }
    const double c = dt / dtau;
    if constexpr (sizeof(Real) == sizeof(float)) {
        k.c = step_ratio_f32(c);
        k.omc = 1.0f - k.c;   // exact
        k.opc = 1.0f + k.c;   // exact
    } else {
        k.c = (Real)c;
        k.omc = (Real)(1.0 - c);
        k.opc = (Real)(1.0 + c);
    }
    k.sqrt_dt = (Real)sqrt(dt);
    k.dtau = (Real)dtau;
}

inline int64_t quantile_rank(int64_t n, double q)
{
    int64_t k = (int64_t)ceil(q * (double)n);
    if (k < 1) k = 1;
    if (k > n) k = n;
    return k;
}

inline void cva_weights(int ndates, double date_dt, double h, double R, double* w)
{
    for (int i = 0; i < ndates; ++i) w[i] = 0.0;
    for (int i = 1; i < ndates; ++i) {
        const double pd = exp(-h * (i - 1) * date_dt) - exp(-h * i * date_dt);
        w[i - 1] += 0.5 * (1.0 - R) * pd;
        w[i] += 0.5 * (1.0 - R) * pd;
    }
}


// Cumulative Simpson integral of y on a uniform grid (the rule of hjm_hjm_drift).
inline void cumulative_simpson(const double* y, double h, double* out)
{
    out[0] = 0.0;
    for (int j = 1; j < kTenors; ++j) {
        double s = 0.0;
        if (j == 1) {
            s = h / 12.0 * (5.0 * y[0] + 8.0 * y[1] - y[2]);
        } else {
            const int even_end = (j % 2 == 0) ? j : j - 3;
            for (int i = 0; i < even_end; i += 2) s += h / 3.0 * (y[i] + 4.0 * y[i + 1] + y[i + 2]);
            if (j % 2 == 1) s += 3.0 * h / 8.0 * (y[j - 3] + 3.0 * y[j - 2] + 3.0 * y[j - 1] + y[j]);
        }
        out[j] = s;
    }
}

// out += (d m / d sigma)^T g for m_j = sum_k sigma_kj I_kj, I_k = W sigma_k:
// d m_j / d sigma_kl = delta_jl I_kj + sigma_kj W_jl.
// Sums over the S steps of a date block of the adjoint's reverse recursion
// lambda(k + 1) = T^T lambda(k) + mu, lambda(0) = lambda_e (T = I + c Dfwd with the
// backward difference on the last tenor, c as the FP32 step takes it):
//   sum_{k<S} lambda(k) = A lambda_e + B mu,  A = sum_{k<S} (T^T)^k,
//   B = sum_{r<S-1} (S-1-r) (T^T)^r   (row-major [51][51]).
inline void adjoint_block_sums(double c, int S, double* A, double* B)
{
    const int n = kTenors;
    std::vector<double> Pk((size_t)n * n, 0.0), next((size_t)n * n);
    for (int j = 0; j < n; ++j) Pk[(size_t)j * n + j] = 1.0;
    for (size_t e = 0; e < (size_t)n * n; ++e) A[e] = B[e] = 0.0;
    for (int k = 0; k < S; ++k) {
        for (size_t e = 0; e < (size_t)n * n; ++e) {
            A[e] += Pk[e];
            if (k < S - 1) B[e] += (double)(S - 1 - k) * Pk[e];
        }
        // next = T^T Pk: row j of T^T holds T[i][j] over i
        for (int j = 0; j < n; ++j)
            for (int m = 0; m < n; ++m) {
                double v;
                if (j == n - 1) v = (1.0 + c) * Pk[(size_t)j * n + m] + c * Pk[(size_t)(j - 1) * n + m];
                else if (j == n - 2) v = (1.0 - c) * Pk[(size_t)j * n + m] + c * Pk[(size_t)(j - 1) * n + m] - c * Pk[(size_t)(j + 1) * n + m];
                else if (j == 0) v = (1.0 - c) * Pk[(size_t)j * n + m];
                else v = (1.0 - c) * Pk[(size_t)j * n + m] + c * Pk[(size_t)(j - 1) * n + m];
                next[(size_t)j * n + m] = v;
            }
        Pk.swap(next);
    }
}

inline void drift_vjp(const double* vols, double h, const double* g, double* out)
{
    double W[kTenors][kTenors];  // W[l][j] = weight of y_l in the integral to node j
    for (int l = 0; l < kTenors; ++l) {
        double e[kTenors] = {0.0};
        e[l] = 1.0;
        cumulative_simpson(e, h, W[l]);
    }
    for (int k = 0; k < kFactors; ++k) {
        const double* sk = vols + k * kTenors;
        double I[kTenors];
        cumulative_simpson(sk, h, I);
        for (int l = 0; l < kTenors; ++l) {
            double a = g[l] * I[l];
            for (int j = 0; j < kTenors; ++j) a += g[j] * sk[j] * W[l][j];
            out[k * kTenors + l] += a;
        }
    }
}

}  // namespace hjm_host
