/*
 * hjm_cva.h -- C ABI of the MI355X HJM Monte Carlo exposure library (libhjm_cva.so).
 *
 * Plain C types only (pointers to host memory, sizes, POD structs); every function
 * returns an hjm_status code instead of printing and continuing as the reference does
 * (simulation_gpu.cu:71-106). All arrays are caller-owned host memory. Calls are
 * synchronous and re-entrant: no process-global market data (the reference keeps it
 * in __constant__ symbols, simulation_gpu.cu:109-119), per-call or per-plan buffers.
 *
 * Grid: 51 Musiela tenors tau_j = j * dtau (j = 0..50), 3 volatility factors,
 * Euler step dt; exposure dates every dtau. The reference's public entry points
 * (simulation_gpu.h:10,12) are re-exposed with their exact C++ signatures by
 * include/simulation_gpu.h on top of this ABI.
 */
#ifndef HJM_CVA_H
#define HJM_CVA_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HJM_TENORS 51
#define HJM_FACTORS 3
#define HJM_MAX_SWAPS 4096
#define HJM_MAX_SETS 512

enum hjm_status_code {
    HJM_OK = 0,
    HJM_ERR_NULL = 1,        /* a required pointer is NULL */
    HJM_ERR_ARG = 2,         /* invalid scalar argument (paths <= 0, dt <= 0, ...) */
    HJM_ERR_GRID = 3,        /* dt does not divide dtau, horizon not a whole number of
                                blocks, or a date beyond the 51-point tenor grid */
    HJM_ERR_DEVICE = 4,      /* no HIP device, or a HIP runtime error */
    HJM_ERR_ALLOC = 5,       /* device memory allocation failed */
    HJM_ERR_LAUNCH = 6,      /* kernel launch or execution failed */
    HJM_ERR_UNSUPPORTED = 7, /* option combination not implemented */
    HJM_ERR_LIMIT = 8        /* size beyond a fixed limit of the call
                                (hjm_simulate_reference_noalloc) */
};

enum hjm_precision {
    HJM_FP64 = 0,       /* FP64 evolution and integrals */
    HJM_FP32_KAHAN = 1  /* FP32 evolution, Kahan-compensated FP32 integrals; FP64 stats */
};

enum hjm_rng_mode {
    HJM_RNG_PHILOX = 0, /* in-kernel Philox4x32-10 + Box-Muller, counter (step, path) */
    HJM_RNG_HOST = 1    /* caller-supplied normals, layout [path][step][3] */
};

enum hjm_path_kernel {
    HJM_KERNEL_AUTO = 0,   /* matrix cores for FP32 + in-kernel RNG when dt = dtau / 50;
                              FP64 on the vector kernel */
    HJM_KERNEL_VECTOR = 1, /* one path per lane, every step on the vector ALUs */
    HJM_KERNEL_MATRIX = 2  /* 50-step blocks as matrix-core products: FP32 as bf16 split
                              products, FP64 with v_mfma_f64; in-kernel RNG and
                              dt = dtau / 50 only (else HJM_ERR_UNSUPPORTED) */
};

/* How the adjoint's reverse sweep gets the forward pass's shocks. */
enum hjm_adjoint_normals {
    HJM_ADJ_NORMALS_AUTO = 0,       /* the matrix-core forward keeps them (paths x steps x 3
                                       FP32) when that much device memory is free; otherwise,
                                       and after a VALU forward, the sweep regenerates them */
    HJM_ADJ_NORMALS_REGENERATE = 1  /* always regenerate (no stored-shock buffer) */
};

/* Market data (simulation_gpu.h:10 arguments accrual, spot_rates, drift, volatilities). */
typedef struct {
    const double* spot_rates;   /* [51] initial forward curve f(0, tau_j) */
    const double* drifts;       /* [51] HJM drift m(tau_j) */
    const double* volatilities; /* [3*51] factor-major loadings sigma_k(tau_j) */
    const double* accrual;      /* [51] year fraction of the cash flow at date j */
} hjm_market;

/* The fields of InterestRateSwap<real> that the reference reads (product.h:4-17). */
typedef struct {
    double notional;
    double fixed_rate; /* K */
    double expiry;     /* years; dates at dtau, 2 dtau, ..., expiry */
    double dtau;       /* tenor spacing = payment period (0.5) */
} hjm_swap;

typedef struct {
    int64_t paths;          /* Monte Carlo paths (scenarios) in this call */
    int64_t path_offset;    /* global id of the first path: shards of one run share a
                               seed and take disjoint offsets */
    double dt;              /* Euler step (years) */
    int32_t precision;      /* enum hjm_precision */
    int32_t rng;            /* enum hjm_rng_mode */
    uint64_t seed;          /* Philox key (HJM_RNG_PHILOX) */
    const double* host_normals; /* HJM_RNG_HOST: [paths][expiry/dt][3] N(0,1) variates */
    double hazard_rate;     /* flat counterparty hazard rate for CVA */
    double recovery;        /* recovery rate for CVA */
    double pfe_quantile;    /* e.g. 0.975; <= 0 skips the PFE selection */
    int32_t device;         /* HIP device ordinal */
    int32_t path_kernel;    /* enum hjm_path_kernel (reference and portfolio path kernels) */
    int32_t adjoint_normals; /* enum hjm_adjoint_normals (the CVA Greeks calls only) */
} hjm_config;

/* Outputs; every pointer may be NULL (then that output is not produced / copied). */
typedef struct {
    double* expected_exposure; /* [51] EE(t_j), discounted as in the reference */
    double* pfe;               /* [51] PFE quantile of the exposure at t_j */
    double* cva;               /* [1] (1-R) sum_j w_j EE_j */
    double* cva_stderr;        /* [1] Monte Carlo standard error of cva */
    double* sums;              /* [53] raw FP64 sums: sum_p E_pj (51), sum_p CVA_p,
                                  sum_p CVA_p^2 -- shards combine these */
    double* num_x;             /* [paths][51] numeraire forward (slot j), 0 if unused */
    double* num_y;             /* [paths][51] numeraire discount factor (slot j) */
    double* exposures;         /* [paths][51] per-path exposure E_pj */
    double* timings_ms;        /* [5] path kernel, exposure kernel, reduction,
                                  PFE selection, all device work (HIP events) */
} hjm_results;

const char* hjm_status_string(int status);
int hjm_device_count(int* count);

/* CVA weights w_j = (1-R) * trapezoid share of the default probability of
 * [t_{j-1}, t_j], S(t) = exp(-h t), t_j = j * date_dt, j = 0..ndates-1. */
int hjm_cva_weights(int ndates, double date_dt, double hazard_rate, double recovery,
                    double* weights);

/* HJM no-arbitrage drift m(tau_j) = sum_k sigma_k(tau_j) int_0^tau_j sigma_k, cumulative
 * Simpson on the tenor grid (reproduces the reference's drift table,
 * simulation_gpu_testing.cpp:66-71, from its volatilities, :47-63). */
int hjm_hjm_drift(const double* volatilities, double dtau, double* drifts);

/* ---- reference semantics: single payer IRS, diagonal numeraire ------------------
 * Path generation and exposure of simulation_cpu.h:188-296 / simulation_gpu.cu:471-587
 * (block 0 frozen at the initial curve, numeraire slot i = tenor i after block i-1,
 * exposure = positive part of the sum of later discounted cash flows).            */
typedef struct hjm_plan_s* hjm_plan;
int hjm_plan_create(hjm_plan* plan, const hjm_market* market, const hjm_swap* swap,
                    const hjm_config* config);
int hjm_plan_execute(hjm_plan plan, hjm_results* results);
void hjm_plan_destroy(hjm_plan plan);
int hjm_simulate_reference(const hjm_market* market, const hjm_swap* swap,
                           const hjm_config* config, hjm_results* results);

/* No counterpart in the reference. For callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the
 * work of hjm_simulate_reference with the one-path-per-lane path kernel (FP64, or FP32 +
 * compensated) and in-kernel Philox, enqueued on the caller's stream without any
 * allocation or synchronisation. Every intermediate buffer lives in the library's static
 * device scratch array (26.8 MB per device, held until the process ends).
 *   d_sums  device pointer to 53 doubles: the raw sums of hjm_results.sums (EE numerators
 *           sum_p E_pj for j = 0..50, sum_p CVA_p, sum_p CVA_p^2), written by the enqueued
 *           work; EE = sums / paths, and shards with disjoint path_offset add up.
 *   stream  hipStream_t on config->device (NULL: the null stream).
 * Limits: config->paths <= HJM_NOALLOC_MAX_PATHS per call (HJM_ERR_LIMIT beyond: run
 * larger jobs as several calls with path_offset and add the sums); rng HJM_RNG_PHILOX
 * and path_kernel AUTO or VECTOR (HJM_ERR_UNSUPPORTED otherwise); no PFE (pfe_quantile is
 * ignored). All calls on one device share the scratch array: calls that overlap in time
 * on one device (different streams or threads) race; serialise them on one stream. */
#define HJM_NOALLOC_MAX_PATHS 32768
int hjm_simulate_reference_noalloc(const hjm_market* market, const hjm_swap* swap,
                                   const hjm_config* config, double* d_sums, void* stream);

/* CVA sensitivities by the pathwise adjoint (no counterpart in the reference; its
 * roadmap names them, doc/TODO:6,14). Forward pass as hjm_simulate_reference (use
 * HJM_FP32_KAHAN for the FP32 forward), then a backward sweep over the Euler steps per
 * path, with FP32 adjoint state and FP64 accumulation of the Greeks across paths.     */
typedef struct {
    double* cva;          /* [1] */
    double* cva_stderr;   /* [1] */
    double* d_spot;       /* [51]  dCVA / d f(0, tau_j) */
    double* d_vol;        /* [153] dCVA / d sigma_k(tau_j), drift table held fixed */
    double* d_drift;      /* [51]  dCVA / d m(tau_j) */
    double* d_vol_total;  /* [153] d_vol + (d m / d sigma)^T d_drift, m the cumulative-
                             Simpson HJM drift of the loadings (hjm_hjm_drift) */
    double* timings_ms;   /* [3] forward (all kernels), adjoint (all kernels), total */
    int64_t* stored_normals_bytes; /* [1] device bytes of the shocks the matrix-core forward
                                      kept for the sweep; 0 when the sweep regenerated them */
} hjm_greeks;
int hjm_cva_greeks_reference(const hjm_market* market, const hjm_swap* swap,
                             const hjm_config* config, hjm_greeks* greeks);

/* The reference's design on the GPU, kept as the "plain FP64 GPU form" baseline:
 * all normals materialised in HBM first, one workgroup per path with one tenor per
 * thread, curve in LDS, one launch per 50-step block (simulation_gpu.cu:471-669),
 * exposure per workgroup (:803-858). Same outputs as hjm_simulate_reference. */
int hjm_simulate_reference_blockwise(const hjm_market* market, const hjm_swap* swap,
                                     const hjm_config* config, hjm_results* results);

/* ---- portfolio semantics (batched swaps over netting sets) --------------------------
 * Successor of the reference's single-swap product (product.h:4-17), whose diagonal
 * numeraire (simulation_cpu.h:248-250,279-295) cannot express maturities past the
 * 25y tenor grid. Dates t_e = e dtau, e = 0..horizon/dtau (<= 60), evolved from t = 0;
 * bank-account numeraire B(t_e) = exp(dt sum_{n < n_e} f(t_n, 0)); bond prices
 * P(t_e, t_e + k dtau) = exp(-trapezoid integral of the simulated curve), flat beyond
 * 25y; swap value sign N (1 - P_n - K dtau sum_{k<=n} P_k) over its remaining n periods;
 * exposure per netting set max(sum of its swaps, 0) / B(t_e).                      */
typedef struct {
    int32_t nswaps;
    int32_t nsets;
    const int32_t* periods;       /* [nswaps] maturity in dtau periods, 1..horizon/dtau */
    const double* sign_notional;  /* [nswaps] +notional for payer, -notional for receiver */
    const double* fixed_rate;     /* [nswaps] */
    const int32_t* netting_set;   /* [nswaps] in 0..nsets-1 */
    const double* hazard_rate;    /* [nsets] flat hazard rate of each counterparty */
    double recovery;
    double horizon;               /* years */
    double dtau;                  /* date spacing = payment period (0.5) */
} hjm_portfolio;

/* Launch shapes of the portfolio plan's kernels (0 in a field: its default). The VALU
 * path kernel (FP64, host normals, or path_kernel = VECTOR) uses block_size and
 * blocks_per_cu; the matrix-core FP32 path kernel path_tiles, path_waves and
 * blocks_per_cu (0 or 1: its block map fills 84 KB of LDS, one workgroup per CU); the
 * matrix-core valuation value_tiles, value_waves and value_stage. The FP64 matrix-core
 * path kernel (FP64 with path_kernel = MATRIX) has a fixed shape. Results do not depend
 * on the shapes. hjm_portfolio_autotune measures them. */
typedef struct {
    int32_t block_size;     /* threads per workgroup of the VALU path kernel: 64, 128 or 256 */
    int32_t blocks_per_cu;  /* persistent path kernel: grid = blocks_per_cu x CUs workgroups
                               looping over the groups of paths; 0 = one workgroup per group */
    int32_t use_graph;      /* capture all passes in one hipGraph, replay it per execute */
    int32_t fuse_statistics; /* FP32 plans: 1 = value and reduce in one pass that keeps the
                                exposures out of HBM (EE, per-path CVA moments, the PFE's
                                bracket counting); 0 = valuation writes the exposures, a
                                statistics pass reads them */
    int32_t pfe_stage_slots; /* fused pass: PFE candidates a workgroup stages per netting
                                set and date (1..32; 0: 32); a row with more is valued again,
                                whole, for its selection (exact either way) */
    int32_t path_tiles;     /* matrix-core path kernel: N-tiles of 16 paths per wave, and */
    int32_t path_waves;     /* waves per workgroup: {2, 8} (default), {2, 12}, {1, 12}, {1, 16} */
    int32_t value_tiles;    /* matrix-core valuation: N-tiles per wave, 2 or 4 (default) */
    int32_t value_waves;    /* waves per workgroup, 4 (default) or 8 */
    int32_t value_stage;    /* M-tiles (16 sets) of -gamma staged in LDS at a time, 1..8
                               (default 8) */
} hjm_tuning;

typedef struct {
    double* expected_exposure;  /* [nsets][ndates] */
    double* pfe;                /* [nsets][ndates] */
    double* cva;                /* [nsets] (1-R) sum_e w_ge EE_ge with the set's hazard */
    double* cva_stderr;         /* [nsets] */
    double* exposures;          /* [paths][nsets][ndates] per-path exposures, or NULL */
    double* timings_ms;         /* [5] path kernel, valuation, statistics, PFE, total */
} hjm_portfolio_results;

typedef struct hjm_pplan_s* hjm_portfolio_plan;
int hjm_portfolio_plan_create(hjm_portfolio_plan* plan, const hjm_market* market,
                              const hjm_portfolio* portfolio, const hjm_config* config,
                              const hjm_tuning* tuning /* NULL: defaults */);
int hjm_portfolio_plan_execute(hjm_portfolio_plan plan, hjm_portfolio_results* results);
void hjm_portfolio_plan_destroy(hjm_portfolio_plan plan);
int hjm_simulate_portfolio(const hjm_market* market, const hjm_portfolio* portfolio,
                           const hjm_config* config, const hjm_tuning* tuning,
                           hjm_portfolio_results* results);
/* Total portfolio CVA (sum over netting sets) and its pathwise-adjoint sensitivities to
 * the 51 initial forwards, 153 loadings and 51 drifts (hjm_greeks; cva_stderr = NaN:
 * per-set standard errors come from hjm_simulate_portfolio). Forward pass as
 * hjm_simulate_portfolio without PFE; the curve checkpoints at every date seed the
 * backward sweep. timings_ms: forward, adjoint, total. */
int hjm_portfolio_cva_greeks(const hjm_market* market, const hjm_portfolio* portfolio,
                             const hjm_config* config, const hjm_tuning* tuning,
                             hjm_greeks* greeks);
/* Measures launch shapes on the caller's problem and writes the fastest into *best
 * (use_graph set) and the eager time of one execute with it into *best_ms. Matrix-core
 * FP32 plans: the path kernel over its four shapes x {one workgroup per group, persistent
 * one per CU}, the valuation over tiles x waves x LDS stage, then the exposure pass
 * against the fused valuation-statistics pass. VALU path kernels: block sizes {64, 128,
 * 256} x persistent grids {0, 1, 2, 4, 8 workgroups per CU} on a 2-date horizon. FP64
 * plans on the matrix cores (path_kernel = MATRIX) have no shape to choose: defaults. */
int hjm_portfolio_autotune(const hjm_market* market, const hjm_portfolio* portfolio,
                           const hjm_config* config, hjm_tuning* best, double* best_ms);

/* ---- drop-in for simulation_gpu.h:10,12 (calculateExposureMultiGPU) --------------
 * EE only; scenarios are split into num_gpus contiguous shards run on devices
 * g % device_count; in-kernel Philox keyed by seed 1234 (the reference's seed,
 * simulation_gpu.cu:962); `scale` is unused as in the reference. The double overload
 * runs HJM_FP64, the float overload HJM_FP32_KAHAN.                                */
int hjm_calculate_exposure_multi_gpu_f64(double* expected_exposure, double notional, double K,
                                         double expiry, double dtau, const double* accrual,
                                         const double* spot_rates, const double* drift,
                                         const double* volatilities, double scale,
                                         int num_gpus, int scenarios, double dt);
int hjm_calculate_exposure_multi_gpu_f32(float* expected_exposure, float notional, float K,
                                         float expiry, float dtau, const float* accrual,
                                         const float* spot_rates, const float* drift,
                                         const float* volatilities, float scale,
                                         int num_gpus, int scenarios, float dt);

/* The affine 50-step block map of a matrix-core path kernel as a dense FP32 matrix with
 * 224 columns in the kernel's K order (2 chunks of 32 for the previous curve, 5 for the
 * block's shocks and the drift constant), for inspection and tests. Host only.
 * portfolio = 0: reference semantics, [102][224] (rows: curve after the block, then the
 * block's discount-integral increments); portfolio = 1: [52][224] (rows: curve after the
 * block, then the short rate's left-point sum over the block's 50 steps). */
int hjm_block_map(const double* drifts, const double* volatilities, double dt, double dtau,
                  int32_t portfolio, float* out);
/* The FP64 matrix-core kernels' block map as they read it, for inspection and tests
 * (host only): the A operands of v_mfma_f64_16x16x4f64, out[m][j][lane] = W[16 m + lane %
 * 16][column (j, lane / 16)] in FP64 for M-tiles m < mtiles and instructions j < 52 (13
 * curve instructions taking row 16 (j / 4) + 4 (j % 4) + g of the previous block, 39
 * shock values v = j - 13: factor v % 3 of step 4 (v / 3) + g, the constant 1 at step 50
 * factor 0). portfolio = 0: reference map, mtiles 1..7; 1: portfolio map, mtiles 1..4.
 * out holds mtiles x 52 x 64 doubles. */
int hjm_block_map64(const double* drifts, const double* volatilities, double dt, double dtau,
                    int32_t portfolio, int32_t mtiles, double* out);

/* ---- calibration upstream of the path (no code in the reference; its roadmap,
 * doc/TODO:16, names the volatility and drift calibration of the HJM model) ----------
 * Batched rolling-window PCA of historical forward curves: per window the annualised
 * sample covariance of the daily curve differences, its eigendecomposition (GPU, FP64,
 * cyclic Jacobi), loadings sigma_k = sqrt(lambda_k) v_k of the 3 leading factors (sign:
 * components sum > 0), optionally least-squares polynomials in the tenor, and their HJM
 * drift (hjm_hjm_drift). The outputs feed hjm_market directly. */
typedef struct {
    const double* curves;       /* [ncurves][51] forward curves f(t_d, tau_j), oldest first */
    int64_t ncurves;
    int64_t window;             /* curves per window (>= 3; window - 1 differences) */
    int64_t step;               /* curves between consecutive window starts (>= 1) */
    int32_t nwindows;           /* (nwindows - 1) * step + window <= ncurves */
    int32_t nfactors;           /* 3: the path kernels' factor count */
    double obs_per_year;        /* observations per year, e.g. 252 */
    double dtau;                /* tenor spacing of the curves (drift quadrature) */
    const int32_t* fit_degrees; /* [nfactors] polynomial degree (0..5) in the tenor per
                                   loading, < 0 for the raw loading; NULL: all raw */
    int32_t device;
} hjm_calibration;

typedef struct {
    double* eigenvalues;   /* [nwindows][51] covariance eigenvalues, descending */
    double* volatilities;  /* [nwindows][3*51] loadings, factor-major, fitted where asked */
    double* drifts;        /* [nwindows][51] HJM drift of those loadings */
    double* raw_loadings;  /* [nwindows][3*51] sqrt(lambda_k) v_k before fitting, or NULL */
    double* timings_ms;    /* [1] calibration kernel time, or NULL */
} hjm_calibration_results;
int hjm_calibrate_pca(const hjm_calibration* calibration, hjm_calibration_results* results);

/* ---- normals (for tests and host-supplied runs): the in-kernel generator's output
 * for paths [path_offset, path_offset + paths) and steps [0, nsteps) of a run whose
 * exposure dates are `stride` steps apart (dtau / dt; the generator's streams follow the
 * date blocks), written as [paths][nsteps][3] doubles. Runs on the GPU.              */
int hjm_generate_normals(uint64_t seed, int64_t path_offset, int64_t paths, int32_t nsteps,
                         int32_t stride, int32_t device, double* out);
/* The generator's Box-Muller factors on caller words (no counterpart in the reference;
 * for tests of the transform's error bound): radius[i] = sqrt(-2 ln u) of words[i] as the
 * first word of a pair, cosine[i] / sine[i] = cos / sin(2 pi t) of words[i] as the second,
 * exactly as the path kernels evaluate them. 1 <= n <= 2^28.                          */
int hjm_box_muller_parts(const uint32_t* words, int64_t n, int32_t device, float* radius,
                         float* cosine, float* sine);

/* ---- the PFE selection on caller data (no counterpart in the reference): out[r] = the
 * k-th smallest (1-based) of row r of values [rows][n], all >= 0 (exact, by the GPU
 * selection the PFE uses: a bracket from the row's first entries, one counting pass and
 * a radix selection among the entries inside it for rows of >= 32768 values, MSD radix
 * selection of the whole row otherwise). HJM_ERR_ARG for negative or non-finite values. */
int hjm_select_kth(const double* values, int32_t rows, int64_t n, int64_t k, int32_t device,
                   double* out);

#ifdef __cplusplus
}
#endif
#endif
