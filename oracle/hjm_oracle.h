/*
 * hjm_oracle.h -- CPU oracle for the HJM path-generation + swap-exposure hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the oracle: it is loaded by tests/,
 * by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, never by the product
 * library (cva_hjm_cuda_amd/), which must run on the GPU or fail.
 *
 * It restates, in plain FP64 C, the reference's sequential CPU path
 * (simulation_cpu.h:37-54 SDE, :96-108 dF, :188-262 block path generation,
 * :269-296 exposure, :335-343 expected exposure, driver loop :349-458), with the
 * normal variates injected by the caller instead of drawn from cuRAND. On top of the
 * reference it adds what the reference lacks but the product computes: the normal
 * generator specification (Philox4x32-10 + Box-Muller), the HJM drift precompute,
 * PFE / CVA statistics, the portfolio (standard revaluation) semantics and the
 * pathwise adjoint. Every function documents the check that pins it (see tests/).
 */
#ifndef HJM_ORACLE_H
#define HJM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_TENORS 51
#define ORC_FACTORS 3

/* ---- normal generator (the product's Philox words and uniforms bit for bit; its
 *      Box-Muller transform within the bound stated in hjm_oracle.c) ---------------- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void orc_box_muller(uint32_t a, uint32_t b, float* z0, float* z1);
void orc_stream_call(uint64_t seed, uint32_t path, uint32_t block, uint32_t cls, uint32_t call,
                     float z[4]);
void orc_step_normals(uint64_t seed, uint32_t path, uint32_t step, uint32_t stride, float z[3]);
void orc_fill_normals(uint64_t seed, int64_t path0, int npaths, int nsteps, int stride,
                      double* out);

/* ---- market data ---------------------------------------------------------------- */
void orc_cumulative_simpson(const double* y, int n, double h, double* out);
void orc_hjm_drift(const double* vols, int nfactors, int ntenors, double h, double* drift);
void orc_nelson_siegel(double b0, double b1, double b2, double lam, int ntenors, double h,
                       double* curve);
void orc_poly_vols(int ntenors, double h, double* vols);

/* ---- reference semantics (simulation_cpu.h) -------------------------------------- */
int orc_reference_paths(const double* spot, const double* drift, const double* vols,
                        double expiry, double dtau, double dt, int npaths,
                        const double* normals, double* num_x, double* num_y);
void orc_reference_exposure(const double* num_x, const double* num_y, const double* accrual,
                            double notional, double K, double dtau, int npaths,
                            double* exposures);
void orc_expected_exposure(const double* exposures, int npaths, int ndates, double* ee);
int orc_calculate_exposure_cpu(double* expected_exposure, const double* spot,
                               const double* drift, const double* vols,
                               const double* accrual, double notional, double K,
                               double expiry, double dtau, double dt, int npaths,
                               const double* normals, double* exposures);

/* ---- statistics ------------------------------------------------------------------ */
void orc_cva_weights(int ndates, double date_dt, double hazard, double recovery, double* w);
int64_t orc_quantile_rank(int64_t n, double q);
void orc_stats(const double* exposures, int npaths, int ndates, const double* w, double q,
               double* ee, double* pfe, double* cva, double* cva_se);

/* ---- pathwise adjoint (reference semantics) ---------------------------------------
 * CVA = (1/P) sum_p sum_t w_t max(V_pt, 0) and its derivatives with respect to the
 * initial curve (51), the loadings (153, drift table held fixed) and the drift table
 * (51), by a backward sweep over the Euler steps in FP64. */
int orc_reference_cva_greeks(const double* spot, const double* drift, const double* vols,
                             const double* accrual, double notional, double K, double expiry,
                             double dtau, double dt, int npaths, const double* normals,
                             const double* w, double* cva, double* d_spot, double* d_vol,
                             double* d_drift);
/* Jacobian-transpose product of the cumulative-Simpson HJM drift: adds
 * (d drift / d vols)^T g to out (153 entries). */
void orc_drift_vjp(const double* vols, double h, const double* g, double* out);

/* ---- portfolio semantics (standard revaluation; no counterpart in the reference) ----
 * Dates t_e = e dtau, e = 0..horizon/dtau, from t = 0 without a frozen block. The curve
 * follows the same Euler step; the numeraire is the bank account
 * B(t_e) = exp(dt sum_{n < n_e} f_0(n)); bond prices P(t_e, t_e + k dtau) = exp(-I_k)
 * with I_k the trapezoid integral of the simulated curve (flat beyond the last tenor);
 * swap s (n_s periods of dtau, sign +1 payer / -1 receiver, notional N, fixed K):
 *   V_s(t_e) = sign N (1 - P_{n_s - e} - K dtau sum_{k=1}^{n_s - e} P_k) for e < n_s;
 * netting set value V_g = sum_{s in g} V_s; exposure E_ge = max(V_g, 0) / B(t_e).
 * Outputs exposures [path][set][date] and optionally curves [path][date][51] and
 * log bank account [path][date]. */
int orc_portfolio_exposure(const double* spot, const double* drift, const double* vols,
                           double horizon, double dtau, double dt, int npaths,
                           const double* normals, int nswaps, const int* periods,
                           const double* sign_notional, const double* fixed_rate,
                           const int* set_of, int nsets, double* exposures, double* curves,
                           double* log_bank);

/* Total portfolio CVA sum_g (1-R) sum_e w_ge EE_ge (w from each set's hazard) and its
 * pathwise derivatives w.r.t. the initial curve (51), loadings (153), drift (51). */
int orc_portfolio_cva_greeks(const double* spot, const double* drift, const double* vols,
                             double horizon, double dtau, double dt, int npaths,
                             const double* normals, int nswaps, const int* periods,
                             const double* sign_notional, const double* fixed_rate,
                             const int* set_of, int nsets, const double* weights,
                             double* cva, double* d_spot, double* d_vol, double* d_drift);

/* ---- timing helper for the CPU baseline ----------------------------------------- */
double orc_time_reference_paths(const double* spot, const double* drift, const double* vols,
                                double expiry, double dtau, double dt, int npaths,
                                const double* normals, double* num_x, double* num_y,
                                int nthreads);

#ifdef __cplusplus
}
#endif
#endif
