// hjm_capi.hip -- C ABI of libhjm_cva.so: validation, host precompute, plans, drop-ins.
//
// Host code above the kernels. The reference's host orchestration
// (__calculateExposureMultiGPU, simulation_gpu.cu:955-1241) allocates, generates all
// normals, launches one kernel per step (or per 50-step block), runs a gemv and sums
// per-GPU partials on the host; here a plan holds device buffers and a stream, one
// launch covers every step of every path, and reductions are deterministic FP64.
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <mutex>
#include <vector>

#include "hjm_common.h"
#include "hjm_kernels.h"
#include "hjm_host.h"
#include "hjm_mfma_plan.h"
#include "hjm_select.h"

using namespace hjm;
using namespace hjm_host;

namespace {

// Time grid. The reference truncates: stride = (int)(dtau/dt) (simulation_cpu.h:200,
// simulation_gpu.cu:494), steps = (int)(expiry/dt) (:959). Here both ratios are rounded
// to the nearest integer and must be integral to 1e-6 relative (so a float dt of 0.01f
// still gives 50 steps per block), the horizon must be a whole number of blocks (else
// the reference's last block reads the next path's normals), and the last numeraire
// slot must lie on the 51-point grid (else the reference writes past the curve).
int ref_grid(double expiry, double dtau, double dt, int* stride, int* nsteps, int* ndates)
{
    if (!(dt > 0.0) || !(dtau > 0.0) || !(expiry > 0.0) || !isfinite(expiry)) return HJM_ERR_ARG;
    const double rs = dtau / dt, rS = expiry / dt;
    if (rs > 1e7 || rS > 1e9) return HJM_ERR_GRID;
    const int st = (int)llround(rs);
    const int S = (int)llround(rS);
    if (st <= 0 || fabs(rs - st) > 1e-6 * st || fabs(rS - S) > 1e-6 * S) return HJM_ERR_GRID;
    if (S < st || S % st != 0) return HJM_ERR_GRID;
    const int D = S / st;
    if (D > kTenors - 1) return HJM_ERR_GRID;
    *stride = st;
    *nsteps = S;
    *ndates = D;
    return HJM_OK;
}

}  // namespace

struct hjm_plan_s {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[5] = {};
    int precision = HJM_FP64;
    bool host_rng = false;
    bool do_pfe = false;
    bool use_mfma = false;
    int64_t paths = 0;
    int64_t nblocks = 0;
    int64_t pfe_rank = 1;
    RefStepConsts<double> kd{};
    RefStepConsts<float> kf{};
    RefGrid grid{};
    ExposureParams ep{};
    MfmaGrid mg{};
    uint16_t* d_wfrag = nullptr;  // bf16 parts of the block map (FP32 plans)
    double* d_wfrag64 = nullptr;  // the block map's FP64 fragments (FP64 plans)
    double* d_normals = nullptr;
    double* d_nx = nullptr;
    double* d_ny = nullptr;
    double* d_ex = nullptr;
    double* d_partials = nullptr;
    double* d_sums = nullptr;  // kStatSlots sums, then the kTenors PFE values (d_pfe)
    double* d_pfe = nullptr;
    char* d_selws = nullptr;  // bracket-selection workspace (PFE)
    double* h_res = nullptr;  // pinned host copy of d_sums (one asynchronous copy per run)
};

namespace {

void plan_release(hjm_plan p)
{
    if (!p) return;
    int prev = -1;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(p->device);
    dev_free(p->d_normals);
    dev_free(p->d_wfrag);
    dev_free(p->d_wfrag64);
    dev_free(p->d_nx);
    dev_free(p->d_ny);
    dev_free(p->d_ex);
    dev_free(p->d_partials);
    dev_free(p->d_sums);
    dev_free(p->d_selws);
    if (p->h_res) (void)hipHostFree(p->h_res);
    for (auto& e : p->ev)
        if (e) (void)hipEventDestroy(e);
    if (p->stream) (void)hipStreamDestroy(p->stream);
    if (prev >= 0) (void)hipSetDevice(prev);
    delete p;
}

// Argument checks of a reference-semantics run (plans and the no-allocation entry point)
// and its time grid.
int check_ref_config(const hjm_market* market, const hjm_swap* swap, const hjm_config* cfg,
                     int* stride, int* S, int* D)
{
    HJM_TRY(check_market(market));
    if (!swap || !cfg) return HJM_ERR_NULL;
    if (cfg->paths <= 0 || cfg->path_offset < 0 || cfg->path_offset + cfg->paths > (1ll << 32))
        return HJM_ERR_ARG;
    if (cfg->precision != HJM_FP64 && cfg->precision != HJM_FP32_KAHAN) return HJM_ERR_ARG;
    if (cfg->rng != HJM_RNG_PHILOX && cfg->rng != HJM_RNG_HOST) return HJM_ERR_ARG;
    if (cfg->rng == HJM_RNG_HOST && !cfg->host_normals) return HJM_ERR_NULL;
    if (!isfinite(swap->notional) || !isfinite(swap->fixed_rate)) return HJM_ERR_ARG;
    if (!(cfg->hazard_rate >= 0.0) || !(cfg->recovery >= 0.0 && cfg->recovery <= 1.0))
        return HJM_ERR_ARG;
    if (cfg->pfe_quantile > 1.0) return HJM_ERR_ARG;
    if (cfg->path_kernel < HJM_KERNEL_AUTO || cfg->path_kernel > HJM_KERNEL_MATRIX) return HJM_ERR_ARG;
    if (cfg->adjoint_normals != HJM_ADJ_NORMALS_AUTO && cfg->adjoint_normals != HJM_ADJ_NORMALS_REGENERATE)
        return HJM_ERR_ARG;
    HJM_TRY(ref_grid(swap->expiry, swap->dtau, cfg->dt, stride, S, D));
    if ((int64_t)cfg->paths * kTenors > (1ll << 40)) return HJM_ERR_ARG;
    return HJM_OK;
}

// Kernel parameters of a checked run: step constants, grid, exposure/CVA parameters.
void fill_ref_params(const hjm_market* market, const hjm_swap* swap, const hjm_config* cfg,
                     int stride, int S, int D, RefStepConsts<double>& kd, RefStepConsts<float>& kf,
                     RefGrid& g, ExposureParams& e)
{
    fill_consts(kd, market, cfg->dt, swap->dtau);
    fill_consts(kf, market, cfg->dt, swap->dtau);
    g.stride = stride;
    g.nsteps = S;
    g.ndates = D;
    g.paths = cfg->paths;
    g.path0 = cfg->path_offset;
    g.seed_lo = (uint32_t)cfg->seed;
    g.seed_hi = (uint32_t)(cfg->seed >> 32);
    g.dt = cfg->dt;
    for (int j = 0; j < kTenors; ++j) {
        // block 0 keeps the initial curve for `stride` steps (simulation_cpu.h:212-213, 235)
        double a = 0.0;
        const double f0 = (cfg->precision == HJM_FP64) ? market->spot_rates[j]
                                                       : (double)(float)market->spot_rates[j];
        for (int s = 0; s < stride; ++s) a += f0;
        g.a0[j] = a;
    }
    e.paths = cfg->paths;
    e.ndates = D;
    e.notional = swap->notional;
    e.K = swap->fixed_rate;
    e.m = 1.0 / swap->dtau;
    for (int j = 0; j < kTenors; ++j) e.accrual[j] = market->accrual[j];
    cva_weights(kTenors, swap->dtau, cfg->hazard_rate, cfg->recovery, e.w);
}

int plan_init(hjm_plan p, const hjm_market* market, const hjm_swap* swap, const hjm_config* cfg)
{
    int stride, S, D;
    HJM_TRY(check_ref_config(market, swap, cfg, &stride, &S, &D));
    const bool mfma_ok = cfg->rng == HJM_RNG_PHILOX && stride == kMfmaStride;
    if (cfg->path_kernel == HJM_KERNEL_MATRIX && !mfma_ok) return HJM_ERR_UNSUPPORTED;
    // AUTO: the matrix cores for FP32; FP64 keeps the vector kernel, which evolves step by
    // step in the reference's operation order, unless MATRIX is asked for
    p->use_mfma = mfma_ok && (cfg->path_kernel == HJM_KERNEL_MATRIX ||
                              (cfg->path_kernel == HJM_KERNEL_AUTO && cfg->precision == HJM_FP32_KAHAN));

    p->device = cfg->device;
    p->precision = cfg->precision;
    p->host_rng = cfg->rng == HJM_RNG_HOST;
    p->paths = cfg->paths;
    p->nblocks = (cfg->paths + kBlock - 1) / kBlock;
    p->do_pfe = cfg->pfe_quantile > 0.0;
    p->pfe_rank = quantile_rank(cfg->paths, cfg->pfe_quantile);
    fill_ref_params(market, swap, cfg, stride, S, D, p->kd, p->kf, p->grid, p->ep);
    const RefGrid& g = p->grid;

    DeviceGuard guard;
    HJM_TRY(guard.set(cfg->device));
    HJM_HIP(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking), HJM_ERR_DEVICE);
    for (auto& ev : p->ev) HJM_HIP(hipEventCreate(&ev), HJM_ERR_DEVICE);
    const size_t slots = (size_t)kTenors * cfg->paths;
    HJM_TRY(dev_alloc(&p->d_nx, slots));
    HJM_TRY(dev_alloc(&p->d_ny, slots));
    HJM_TRY(dev_alloc(&p->d_ex, slots));
    // exposure rows past the last date are never written by a run
    HJM_HIP(hipMemset(p->d_ex, 0, slots * sizeof(double)), HJM_ERR_DEVICE);
    HJM_TRY(dev_alloc(&p->d_partials, (size_t)p->nblocks * kStatSlots));
    HJM_TRY(dev_alloc(&p->d_sums, kStatSlots + kTenors));
    p->d_pfe = p->d_sums + kStatSlots;
    HJM_HIP(hipHostMalloc((void**)&p->h_res, sizeof(double) * (kStatSlots + kTenors),
                          hipHostMallocDefault),
            HJM_ERR_ALLOC);
    if (p->do_pfe) {
        const size_t wsb = bracket_workspace_bytes(
            kTenors, bracket_plan(cfg->paths, p->pfe_rank, sizeof(double)), sizeof(double));
        if (wsb) HJM_TRY(dev_alloc(&p->d_selws, wsb));
    }
    // host copies of the block map, kept until the stream is synchronised below
    std::vector<uint16_t> wf;
    std::vector<double> wf64;
    if (p->use_mfma) {
        if (cfg->precision == HJM_FP64) {
            HJM_HIP(mfma64_kernels_prepare(), HJM_ERR_DEVICE);
            build_mfma64_block_map(market->drifts, market->volatilities, cfg->dt, swap->dtau,
                                   ref_mfma_tiles(D), wf64);
            HJM_TRY(dev_alloc(&p->d_wfrag64, wf64.size()));
            HJM_HIP(hipMemcpyAsync(p->d_wfrag64, wf64.data(), wf64.size() * sizeof(double),
                                   hipMemcpyHostToDevice, p->stream),
                    HJM_ERR_DEVICE);
        } else {
            HJM_HIP(mfma_kernels_prepare(), HJM_ERR_DEVICE);
            build_mfma_block_map(market->drifts, market->volatilities, cfg->dt, swap->dtau, wf);
            HJM_TRY(dev_alloc(&p->d_wfrag, wf.size()));
            HJM_HIP(hipMemcpyAsync(p->d_wfrag, wf.data(), wf.size() * sizeof(uint16_t),
                                   hipMemcpyHostToDevice, p->stream),
                    HJM_ERR_DEVICE);
        }
        MfmaGrid& mg = p->mg;
        mg.ndates = D;
        mg.paths = cfg->paths;
        mg.path0 = cfg->path_offset;
        mg.seed_lo = g.seed_lo;
        mg.seed_hi = g.seed_hi;
        mg.dt = cfg->dt;
        for (int j = 0; j < kTenors; ++j) {
            mg.a0[j] = g.a0[j];
            mg.spot[j] = (float)market->spot_rates[j];
            mg.spot64[j] = market->spot_rates[j];
        }
    }
    HJM_HIP(hipMemsetAsync(p->d_nx, 0, slots * sizeof(double), p->stream), HJM_ERR_DEVICE);
    HJM_HIP(hipMemsetAsync(p->d_ny, 0, slots * sizeof(double), p->stream), HJM_ERR_DEVICE);
    if (p->host_rng) {
        const size_t n = (size_t)cfg->paths * S * 3;
        HJM_TRY(dev_alloc(&p->d_normals, n));
        HJM_HIP(hipMemcpyAsync(p->d_normals, cfg->host_normals, n * sizeof(double),
                               hipMemcpyHostToDevice, p->stream),
                HJM_ERR_DEVICE);
    }
    HJM_HIP(hipStreamSynchronize(p->stream), HJM_ERR_DEVICE);
    return HJM_OK;
}

// Enqueue every kernel of one run on the plan's stream (asynchronous).
int plan_launch(hjm_plan p, bool keep_exposures, float* zs = nullptr)
{
    hipStream_t st = p->stream;
    HJM_HIP(hipEventRecord(p->ev[0], st), HJM_ERR_DEVICE);
    const bool f64 = p->precision == HJM_FP64;
    hipError_t err;
    if (p->use_mfma)
        err = f64 ? launch_ref_paths_mfma64(p->mg, p->d_wfrag64, p->d_nx, p->d_ny, zs, st)
                  : launch_ref_paths_mfma(p->mg, p->d_wfrag, p->d_nx, p->d_ny, zs, st);
    else
        err = f64 ? launch_ref_paths_f64(p->kd, p->grid, p->d_normals, p->d_nx, p->d_ny, st)
                  : launch_ref_paths_f32k(p->kf, p->grid, p->d_normals, p->d_nx, p->d_ny, st);
    HJM_HIP(err, HJM_ERR_LAUNCH);
    HJM_HIP(hipEventRecord(p->ev[1], st), HJM_ERR_DEVICE);
    double* ex = (p->do_pfe || keep_exposures) ? p->d_ex : nullptr;
    HJM_HIP(launch_ref_exposure(p->ep, p->d_nx, p->d_ny, ex, p->d_partials, st), HJM_ERR_LAUNCH);
    HJM_HIP(hipEventRecord(p->ev[2], st), HJM_ERR_DEVICE);
    HJM_HIP(launch_column_sum(p->d_partials, p->nblocks, kStatSlots, p->d_sums, st), HJM_ERR_LAUNCH);
    HJM_HIP(hipEventRecord(p->ev[3], st), HJM_ERR_DEVICE);
    if (p->do_pfe)
        HJM_HIP(launch_radix_select(p->d_ex, p->grid.ndates, p->paths, p->pfe_rank, p->d_selws,
                                    p->d_pfe, st),
                HJM_ERR_LAUNCH);
    HJM_HIP(hipEventRecord(p->ev[4], st), HJM_ERR_DEVICE);
    HJM_HIP(hipMemcpyAsync(p->h_res, p->d_sums, sizeof(double) * (kStatSlots + kTenors),
                           hipMemcpyDeviceToHost, st),
            HJM_ERR_DEVICE);
    return HJM_OK;
}

void finish_stats(const double* sums, int64_t P, const double* w, double* ee, double* cva,
                  double* se)
{
    const double alpha = 1.0 / (double)P;
    double e[kTenors];
    double c = 0.0;
    for (int t = 0; t < kTenors; ++t) {
        e[t] = alpha * sums[t];
        c += w[t] * e[t];
    }
    if (ee) memcpy(ee, e, sizeof(e));
    if (cva) *cva = c;
    if (se) {
        const double s1 = sums[kTenors], s2 = sums[kTenors + 1];
        const double var = (P > 1) ? (s2 - s1 * s1 / (double)P) / (double)(P - 1) : 0.0;
        *se = sqrt((var > 0.0 ? var : 0.0) / (double)P);
    }
}

// Wait for the run, copy results to the caller, finish the statistics on the host.
int plan_collect(hjm_plan p, hjm_results* r)
{
    HJM_HIP(hipStreamSynchronize(p->stream), HJM_ERR_LAUNCH);
    const double* sums = p->h_res;
    if (r->sums) memcpy(r->sums, sums, sizeof(double) * kStatSlots);
    finish_stats(sums, p->paths, p->ep.w, r->expected_exposure, r->cva, r->cva_stderr);
    if (r->pfe) {
        for (int t = 0; t < kTenors; ++t) r->pfe[t] = 0.0;
        if (p->do_pfe) memcpy(r->pfe, p->h_res + kStatSlots, sizeof(double) * p->grid.ndates);
    }
    const int64_t P = p->paths;
    auto transpose_back = [&](const double* dsrc, double* dst) -> int {
        std::vector<double> tmp((size_t)kTenors * P);
        if (hipMemcpy(tmp.data(), dsrc, tmp.size() * sizeof(double), hipMemcpyDeviceToHost) !=
            hipSuccess)
            return HJM_ERR_DEVICE;
        for (int64_t i = 0; i < P; ++i)
            for (int t = 0; t < kTenors; ++t) dst[i * kTenors + t] = tmp[(size_t)t * P + i];
        return HJM_OK;
    };
    if (r->num_x) HJM_TRY(transpose_back(p->d_nx, r->num_x));
    if (r->num_y) HJM_TRY(transpose_back(p->d_ny, r->num_y));
    if (r->exposures) HJM_TRY(transpose_back(p->d_ex, r->exposures));
    if (r->timings_ms) {
        float ms = 0.f;
        for (int i = 0; i < 4; ++i) {
            HJM_HIP(hipEventElapsedTime(&ms, p->ev[i], p->ev[i + 1]), HJM_ERR_DEVICE);
            r->timings_ms[i] = ms;
        }
        HJM_HIP(hipEventElapsedTime(&ms, p->ev[0], p->ev[4]), HJM_ERR_DEVICE);
        r->timings_ms[4] = ms;
    }
    return HJM_OK;
}

}  // namespace

extern "C" {

const char* hjm_status_string(int status)
{
    switch (status) {
    case HJM_OK: return "ok";
    case HJM_ERR_NULL: return "required pointer is NULL";
    case HJM_ERR_ARG: return "invalid argument";
    case HJM_ERR_GRID: return "inconsistent time grid (dt must divide dtau, whole blocks, "
                              "dates within the 51-point tenor grid)";
    case HJM_ERR_DEVICE: return "HIP device error or no device";
    case HJM_ERR_ALLOC: return "device allocation failed";
    case HJM_ERR_LAUNCH: return "kernel launch or execution failed";
    case HJM_ERR_UNSUPPORTED: return "unsupported option combination";
    case HJM_ERR_LIMIT: return "size beyond a fixed limit of the call (the no-allocation "
                               "entry point runs at most HJM_NOALLOC_MAX_PATHS paths)";
    default: return "unknown status";
    }
}

int hjm_device_count(int* count)
{
    if (!count) return HJM_ERR_NULL;
    *count = 0;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        return HJM_ERR_DEVICE;
    }
    *count = n;
    return HJM_OK;
}

int hjm_cva_weights(int ndates, double date_dt, double hazard_rate, double recovery,
                    double* weights)
{
    if (!weights) return HJM_ERR_NULL;
    if (ndates <= 0 || !(date_dt > 0.0) || !(hazard_rate >= 0.0) ||
        !(recovery >= 0.0 && recovery <= 1.0))
        return HJM_ERR_ARG;
    cva_weights(ndates, date_dt, hazard_rate, recovery, weights);
    return HJM_OK;
}

// Cumulative Simpson on a uniform grid: composite 1/3 rule to even nodes, 3/8 rule on
// the last three intervals for odd nodes, the 3-point quadratic on the first interval.
int hjm_hjm_drift(const double* vols, double dtau, double* drifts)
{
    if (!vols || !drifts) return HJM_ERR_NULL;
    if (!(dtau > 0.0)) return HJM_ERR_ARG;
    const double h = dtau;
    for (int j = 0; j < kTenors; ++j) drifts[j] = 0.0;
    for (int k = 0; k < kFactors; ++k) {
        const double* y = vols + k * kTenors;
        for (int j = 1; j < kTenors; ++j) {
            double s = 0.0;
            if (j == 1) {
                s = h / 12.0 * (5.0 * y[0] + 8.0 * y[1] - y[2]);
            } else {
                const int even_end = (j % 2 == 0) ? j : j - 3;
                for (int i = 0; i < even_end; i += 2) s += h / 3.0 * (y[i] + 4.0 * y[i + 1] + y[i + 2]);
                if (j % 2 == 1) s += 3.0 * h / 8.0 * (y[j - 3] + 3.0 * y[j - 2] + 3.0 * y[j - 1] + y[j]);
            }
            drifts[j] += y[j] * s;
        }
    }
    return HJM_OK;
}

int hjm_plan_create(hjm_plan* out, const hjm_market* market, const hjm_swap* swap,
                    const hjm_config* config)
{
    if (!out) return HJM_ERR_NULL;
    *out = nullptr;
    hjm_plan p = new (std::nothrow) hjm_plan_s();
    if (!p) return HJM_ERR_ALLOC;
    const int rc = plan_init(p, market, swap, config);
    if (rc != HJM_OK) {
        plan_release(p);
        return rc;
    }
    *out = p;
    return HJM_OK;
}

int hjm_plan_execute(hjm_plan plan, hjm_results* results)
{
    if (!plan || !results) return HJM_ERR_NULL;
    DeviceGuard guard;
    HJM_TRY(guard.set(plan->device));
    HJM_TRY(plan_launch(plan, results->exposures != nullptr));
    return plan_collect(plan, results);
}

void hjm_plan_destroy(hjm_plan plan) { plan_release(plan); }

int hjm_simulate_reference(const hjm_market* market, const hjm_swap* swap,
                           const hjm_config* config, hjm_results* results)
{
    if (!results) return HJM_ERR_NULL;
    hjm_plan p = nullptr;
    HJM_TRY(hjm_plan_create(&p, market, swap, config));
    const int rc = hjm_plan_execute(p, results);
    hjm_plan_destroy(p);
    return rc;
}

int hjm_cva_greeks_reference(const hjm_market* market, const hjm_swap* swap,
                             const hjm_config* cfg, hjm_greeks* gr)
{
    if (!gr) return HJM_ERR_NULL;
    hjm_plan p = nullptr;
    HJM_TRY(hjm_plan_create(&p, market, swap, cfg));
    constexpr int kOut = 5 * kTenors;  // g0, g1, g2, sum of lambda_e, d spot
    struct Cleanup {
        hjm_plan p;
        float *gx = nullptr, *gA = nullptr, *zs = nullptr;
        double *part = nullptr, *sums = nullptr, *sap = nullptr;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        ~Cleanup()
        {
            dev_free(gx);
            dev_free(zs);
            dev_free(gA);
            dev_free(part);
            dev_free(sums);
            dev_free(sap);
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
            hjm_plan_destroy(p);
        }
    } c{p};
    DeviceGuard guard;
    HJM_TRY(guard.set(p->device));
    const int64_t P = p->paths, nb = ref_adjoint_blocks(P);
    const int D = p->grid.ndates, S = p->grid.stride;
    const size_t cells = (size_t)kTenors * P;
    HJM_TRY(dev_alloc(&c.gx, cells));
    HJM_TRY(dev_alloc(&c.gA, cells));
    HJM_TRY(dev_alloc(&c.part, (size_t)nb * kOut));
    HJM_TRY(dev_alloc(&c.sums, kOut + kTenors));  // then the per-slot sums of gA
    HJM_TRY(dev_alloc(&c.sap, (size_t)p->nblocks * kTenors));
    HJM_HIP(hipEventCreate(&c.e0), HJM_ERR_DEVICE);
    HJM_HIP(hipEventCreate(&c.e1), HJM_ERR_DEVICE);
    // the matrix-core forward keeps its normals for the reverse sweep (one FP32 per factor
    // per evolved step and path); without the memory the sweep regenerates them
    if (p->use_mfma && cfg->adjoint_normals == HJM_ADJ_NORMALS_AUTO &&
        dev_alloc(&c.zs, (size_t)p->grid.nsteps * 3 * (size_t)P) != HJM_OK)
        c.zs = nullptr;
    hipStream_t st = p->stream;
    HJM_TRY(plan_launch(p, false, c.zs));
    HJM_HIP(hipEventRecord(c.e0, st), HJM_ERR_DEVICE);
    const double dt = p->grid.dt;
    const double cc = dt / swap->dtau;
    HJM_HIP(launch_ref_adjoint_seed(p->ep, dt, p->d_nx, p->d_ny, c.gx, c.gA, c.sap, st),
            HJM_ERR_LAUNCH);
    // gA summed over the paths at every slot (the drift gradient's mu terms)
    HJM_HIP(launch_column_sum(c.sap, p->nblocks, kTenors, c.sums + kOut, st), HJM_ERR_LAUNCH);
    HJM_HIP(launch_ref_adjoint(p->grid, cc, sqrt(dt), p->d_normals, c.zs, c.gx, c.gA, c.part, st),
            HJM_ERR_LAUNCH);
    HJM_HIP(launch_column_sum(c.part, nb, kOut, c.sums, st), HJM_ERR_LAUNCH);
    HJM_HIP(hipEventRecord(c.e1, st), HJM_ERR_DEVICE);
    hjm_results r{};
    r.cva = gr->cva;
    r.cva_stderr = gr->cva_stderr;
    HJM_TRY(plan_collect(p, &r));
    double sums[kOut + kTenors];
    HJM_HIP(hipMemcpy(sums, c.sums, sizeof(sums), hipMemcpyDeviceToHost), HJM_ERR_DEVICE);
    const double inv = 1.0 / (double)P;
    double dv[kFactors * kTenors], dm[kTenors];
    for (int k = 0; k < kFactors * kTenors; ++k) dv[k] = sums[k] * inv;
    // drift: dt (A Lambda + B M) / P, Lambda = sum over blocks and paths of lambda_e, M_s =
    // (s - 1) x (sum over paths of gA at slot s): the injection at slot s is live in the
    // s - 1 blocks 1 .. s - 1
    {
        std::vector<double> A((size_t)kTenors * kTenors), B((size_t)kTenors * kTenors);
        adjoint_block_sums((double)step_ratio_f32(cc), S, A.data(), B.data());
        double M[kTenors] = {0.0};
        for (int slot = 2; slot <= D; ++slot) M[slot] = (double)(slot - 1) * sums[kOut + slot];
        const double* L = sums + 3 * kTenors;
        for (int j = 0; j < kTenors; ++j) {
            double a = 0.0;
            for (int m = 0; m < kTenors; ++m)
                a += A[(size_t)j * kTenors + m] * L[m] + B[(size_t)j * kTenors + m] * M[m];
            dm[j] = dt * a * inv;
        }
    }
    if (gr->d_spot)
        for (int j = 0; j < kTenors; ++j) gr->d_spot[j] = sums[4 * kTenors + j] * inv;
    if (gr->d_vol) memcpy(gr->d_vol, dv, sizeof(dv));
    if (gr->d_drift) memcpy(gr->d_drift, dm, sizeof(dm));
    if (gr->d_vol_total) {
        memcpy(gr->d_vol_total, dv, sizeof(dv));
        drift_vjp(market->volatilities, swap->dtau, dm, gr->d_vol_total);
    }
    if (gr->stored_normals_bytes)
        *gr->stored_normals_bytes = c.zs ? (int64_t)p->grid.nsteps * 3 * P * (int64_t)sizeof(float) : 0;
    if (gr->timings_ms) {
        float fwd = 0.f, adj = 0.f;
        HJM_HIP(hipEventElapsedTime(&fwd, p->ev[0], p->ev[4]), HJM_ERR_DEVICE);
        HJM_HIP(hipEventElapsedTime(&adj, c.e0, c.e1), HJM_ERR_DEVICE);
        gr->timings_ms[0] = fwd;
        gr->timings_ms[1] = adj;
        gr->timings_ms[2] = fwd + adj;
    }
    return HJM_OK;
}

int hjm_simulate_reference_blockwise(const hjm_market* market, const hjm_swap* swap,
                                     const hjm_config* cfg, hjm_results* r)
{
    if (!r) return HJM_ERR_NULL;
    HJM_TRY(check_market(market));
    if (!swap || !cfg) return HJM_ERR_NULL;
    if (cfg->precision != HJM_FP64) return HJM_ERR_UNSUPPORTED;
    if (cfg->paths <= 0 || cfg->paths > (1ll << 31) - 1 || cfg->path_offset < 0) return HJM_ERR_ARG;
    if (cfg->rng == HJM_RNG_HOST && !cfg->host_normals) return HJM_ERR_NULL;
    int stride, S, D;
    HJM_TRY(ref_grid(swap->expiry, swap->dtau, cfg->dt, &stride, &S, &D));
    DeviceGuard guard;
    HJM_TRY(guard.set(cfg->device));
    const int64_t P = cfg->paths;
    RefStepConsts<double> k{};
    fill_consts(k, market, cfg->dt, swap->dtau);
    RefGrid g{};
    g.stride = stride;
    g.nsteps = S;
    g.ndates = D;
    g.paths = P;
    g.path0 = cfg->path_offset;
    g.dt = cfg->dt;
    ExposureParams e{};
    e.paths = P;
    e.ndates = D;
    e.notional = swap->notional;
    e.K = swap->fixed_rate;
    e.m = 1.0 / swap->dtau;
    for (int j = 0; j < kTenors; ++j) e.accrual[j] = market->accrual[j];
    cva_weights(kTenors, swap->dtau, cfg->hazard_rate, cfg->recovery, e.w);

    struct Bufs {
        double *normals = nullptr, *s0 = nullptr, *s1 = nullptr, *integ = nullptr, *nx = nullptr,
               *ny = nullptr, *ex = nullptr, *sums = nullptr;
        hipStream_t st = nullptr;
        hipEvent_t ev[5] = {};
        ~Bufs()
        {
            dev_free(normals);
            dev_free(s0);
            dev_free(s1);
            dev_free(integ);
            dev_free(nx);
            dev_free(ny);
            dev_free(ex);
            dev_free(sums);
            for (auto& x : ev)
                if (x) (void)hipEventDestroy(x);
            if (st) (void)hipStreamDestroy(st);
        }
    } b;
    const size_t nn = (size_t)P * S * 3, cells = (size_t)P * kTenors;
    HJM_HIP(hipStreamCreateWithFlags(&b.st, hipStreamNonBlocking), HJM_ERR_DEVICE);
    for (auto& x : b.ev) HJM_HIP(hipEventCreate(&x), HJM_ERR_DEVICE);
    HJM_TRY(dev_alloc(&b.normals, nn));
    HJM_TRY(dev_alloc(&b.s0, cells));
    HJM_TRY(dev_alloc(&b.s1, cells));
    HJM_TRY(dev_alloc(&b.integ, cells));
    HJM_TRY(dev_alloc(&b.nx, cells));
    HJM_TRY(dev_alloc(&b.ny, cells));
    HJM_TRY(dev_alloc(&b.ex, cells));
    HJM_TRY(dev_alloc(&b.sums, kTenors));
    HJM_HIP(hipMemsetAsync(b.integ, 0, cells * sizeof(double), b.st), HJM_ERR_DEVICE);
    HJM_HIP(hipMemsetAsync(b.nx, 0, cells * sizeof(double), b.st), HJM_ERR_DEVICE);
    HJM_HIP(hipMemsetAsync(b.ny, 0, cells * sizeof(double), b.st), HJM_ERR_DEVICE);
    HJM_HIP(hipEventRecord(b.ev[0], b.st), HJM_ERR_DEVICE);
    if (cfg->rng == HJM_RNG_HOST)
        HJM_HIP(hipMemcpyAsync(b.normals, cfg->host_normals, nn * sizeof(double),
                               hipMemcpyHostToDevice, b.st),
                HJM_ERR_DEVICE);
    else
        HJM_HIP(launch_normals(cfg->seed, cfg->path_offset, P, S, stride, b.normals, b.st),
                HJM_ERR_LAUNCH);
    HJM_HIP(hipEventRecord(b.ev[1], b.st), HJM_ERR_DEVICE);
    double* cur = b.s0;
    double* nxt = b.s1;
    for (int blk = 0; blk < D; ++blk) {
        HJM_HIP(launch_blockwise_paths(k, g, blk, b.normals, cur, nxt, b.integ, b.nx, b.ny, b.st),
                HJM_ERR_LAUNCH);
        double* t = cur;
        cur = nxt;
        nxt = t;
    }
    HJM_HIP(hipEventRecord(b.ev[2], b.st), HJM_ERR_DEVICE);
    HJM_HIP(launch_blockwise_exposure(e, b.nx, b.ny, b.ex, b.st), HJM_ERR_LAUNCH);
    HJM_HIP(hipEventRecord(b.ev[3], b.st), HJM_ERR_DEVICE);
    HJM_HIP(launch_column_mean_gemv(b.ex, P, b.sums, b.st), HJM_ERR_LAUNCH);
    HJM_HIP(hipEventRecord(b.ev[4], b.st), HJM_ERR_DEVICE);
    HJM_HIP(hipStreamSynchronize(b.st), HJM_ERR_LAUNCH);
    double sums[kStatSlots] = {};
    HJM_HIP(hipMemcpy(sums, b.sums, sizeof(double) * kTenors, hipMemcpyDeviceToHost), HJM_ERR_DEVICE);
    finish_stats(sums, P, e.w, r->expected_exposure, r->cva, nullptr);
    if (r->cva_stderr) *r->cva_stderr = NAN;
    if (r->sums) memcpy(r->sums, sums, sizeof(sums));
    if (r->pfe)
        for (int t = 0; t < kTenors; ++t) r->pfe[t] = NAN;
    auto back = [&](const double* d, double* h, bool slot_major) -> int {
        std::vector<double> tmp(cells);
        if (hipMemcpy(tmp.data(), d, cells * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
            return HJM_ERR_DEVICE;
        for (int64_t i = 0; i < P; ++i)
            for (int t = 0; t < kTenors; ++t)
                h[i * kTenors + t] = slot_major ? tmp[(size_t)t * P + i] : tmp[(size_t)i * kTenors + t];
        return HJM_OK;
    };
    if (r->num_x) HJM_TRY(back(b.nx, r->num_x, true));
    if (r->num_y) HJM_TRY(back(b.ny, r->num_y, true));
    if (r->exposures) HJM_TRY(back(b.ex, r->exposures, false));
    if (r->timings_ms) {
        // [0] normals generation, [1] path blocks, [2] exposure, [3] EE gemv, [4] total
        float ms = 0.f;
        for (int i = 0; i < 4; ++i) {
            HJM_HIP(hipEventElapsedTime(&ms, b.ev[i], b.ev[i + 1]), HJM_ERR_DEVICE);
            r->timings_ms[i] = ms;
        }
        HJM_HIP(hipEventElapsedTime(&ms, b.ev[0], b.ev[4]), HJM_ERR_DEVICE);
        r->timings_ms[4] = ms;
    }
    return HJM_OK;
}

int hjm_block_map(const double* drifts, const double* vols, double dt, double dtau,
                  int32_t portfolio, float* out)
{
    if (!drifts || !vols || !out) return HJM_ERR_NULL;
    if (portfolio != 0 && portfolio != 1) return HJM_ERR_ARG;
    if (!(dt > 0.0) || !(dtau > 0.0) || fabs(dtau / dt - kMfmaStride) > 1e-6 * kMfmaStride)
        return HJM_ERR_GRID;
    const std::vector<double> W = mfma_block_map_dense(
        drifts, vols, dt, dtau, portfolio ? MfmaMap::kPortfolio : MfmaMap::kReference);
    for (size_t i = 0; i < W.size(); ++i) out[i] = (float)W[i];
    return HJM_OK;
}

int hjm_block_map64(const double* drifts, const double* vols, double dt, double dtau,
                    int32_t portfolio, int32_t mtiles, double* out)
{
    if (!drifts || !vols || !out) return HJM_ERR_NULL;
    if (portfolio != 0 && portfolio != 1) return HJM_ERR_ARG;
    if (mtiles < 1 || mtiles > (portfolio ? kMfmaPfMTiles : kMfmaMTiles)) return HJM_ERR_ARG;
    if (!(dt > 0.0) || !(dtau > 0.0) || fabs(dtau / dt - kMfmaStride) > 1e-6 * kMfmaStride)
        return HJM_ERR_GRID;
    std::vector<double> frag;
    build_mfma64_block_map(drifts, vols, dt, dtau, mtiles, frag,
                           portfolio ? MfmaMap::kPortfolio : MfmaMap::kReference);
    memcpy(out, frag.data(), frag.size() * sizeof(double));
    return HJM_OK;
}

int hjm_generate_normals(uint64_t seed, int64_t path_offset, int64_t paths, int32_t nsteps,
                         int32_t stride, int32_t device, double* out)
{
    if (!out) return HJM_ERR_NULL;
    if (paths <= 0 || nsteps <= 0 || stride <= 0 || path_offset < 0 ||
        paths * (int64_t)nsteps > (1ll << 31))
        return HJM_ERR_ARG;
    DeviceGuard guard;
    HJM_TRY(guard.set(device));
    double* d = nullptr;
    const size_t n = (size_t)paths * nsteps * 3;
    HJM_TRY(dev_alloc(&d, n));
    int rc = HJM_OK;
    if (launch_normals(seed, path_offset, paths, nsteps, stride, d, nullptr) != hipSuccess)
        rc = HJM_ERR_LAUNCH;
    else if (hipDeviceSynchronize() != hipSuccess)
        rc = HJM_ERR_LAUNCH;
    else if (hipMemcpy(out, d, n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
        rc = HJM_ERR_DEVICE;
    dev_free(d);
    return rc;
}

int hjm_box_muller_parts(const uint32_t* words, int64_t n, int32_t device, float* radius,
                         float* cosine, float* sine)
{
    if (!words || !radius || !cosine || !sine) return HJM_ERR_NULL;
    if (n <= 0 || n > (1ll << 28)) return HJM_ERR_ARG;
    DeviceGuard guard;
    HJM_TRY(guard.set(device));
    uint32_t* dw = nullptr;
    float* d = nullptr;
    HJM_TRY(dev_alloc(&dw, (size_t)n));
    int rc = dev_alloc(&d, 3 * (size_t)n);
    if (rc == HJM_OK) {
        const size_t b = (size_t)n * sizeof(float);
        if (hipMemcpy(dw, words, (size_t)n * sizeof(uint32_t), hipMemcpyHostToDevice) != hipSuccess)
            rc = HJM_ERR_DEVICE;
        else if (launch_box_muller_parts(dw, n, d, d + n, d + 2 * n, nullptr) != hipSuccess ||
                 hipDeviceSynchronize() != hipSuccess)
            rc = HJM_ERR_LAUNCH;
        else if (hipMemcpy(radius, d, b, hipMemcpyDeviceToHost) != hipSuccess ||
                 hipMemcpy(cosine, d + n, b, hipMemcpyDeviceToHost) != hipSuccess ||
                 hipMemcpy(sine, d + 2 * n, b, hipMemcpyDeviceToHost) != hipSuccess)
            rc = HJM_ERR_DEVICE;
        dev_free(d);
    }
    dev_free(dw);
    return rc;
}

int hjm_select_kth(const double* values, int32_t rows, int64_t n, int64_t k, int32_t device,
                   double* out)
{
    if (!values || !out) return HJM_ERR_NULL;
    if (rows <= 0 || n <= 0 || k < 1 || k > n) return HJM_ERR_ARG;
    const size_t cells = (size_t)rows * (size_t)n;
    for (size_t i = 0; i < cells; ++i)
        if (!(values[i] >= 0.0) || !isfinite(values[i])) return HJM_ERR_ARG;
    DeviceGuard guard;
    HJM_TRY(guard.set(device));
    double *d = nullptr, *d_out = nullptr;
    char* ws = nullptr;
    const size_t wsb = bracket_workspace_bytes(rows, bracket_plan(n, k, sizeof(double)), sizeof(double));
    int rc = dev_alloc(&d, cells);
    if (rc == HJM_OK) rc = dev_alloc(&d_out, (size_t)rows);
    if (rc == HJM_OK && wsb) rc = dev_alloc(&ws, wsb);
    if (rc == HJM_OK && hipMemcpy(d, values, cells * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
        rc = HJM_ERR_DEVICE;
    if (rc == HJM_OK && launch_radix_select(d, rows, n, k, ws, d_out, nullptr) != hipSuccess)
        rc = HJM_ERR_LAUNCH;
    if (rc == HJM_OK && hipDeviceSynchronize() != hipSuccess) rc = HJM_ERR_LAUNCH;
    if (rc == HJM_OK && hipMemcpy(out, d_out, rows * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
        rc = HJM_ERR_DEVICE;
    dev_free(d);
    dev_free(d_out);
    dev_free(ws);
    return rc;
}

// simulation_gpu.cu:955-1241 semantics: scenarios split into num_gpus contiguous shards
// (the remainder is spread over the first shards instead of being dropped, :960), one
// shard per device g % device_count, shards enqueued before any is awaited, partial
// sums combined on the host in shard order and divided by the total (:1154-1179).
static int multi_gpu_ee(double* ee, const hjm_market* mk, const hjm_swap* sw, int precision,
                        int num_gpus, int scenarios, double dt)
{
    if (!ee) return HJM_ERR_NULL;
    if (num_gpus <= 0 || scenarios <= 0) return HJM_ERR_ARG;
    if (num_gpus > scenarios) num_gpus = scenarios;
    int ndev = 0;
    HJM_TRY(hjm_device_count(&ndev));
    if (ndev <= 0) return HJM_ERR_DEVICE;
    std::vector<hjm_plan> plans(num_gpus, nullptr);
    int rc = HJM_OK;
    const int64_t base = scenarios / num_gpus, rem = scenarios % num_gpus;
    int64_t lo = 0;
    for (int gi = 0; gi < num_gpus && rc == HJM_OK; ++gi) {
        hjm_config c{};
        c.paths = base + (gi < rem ? 1 : 0);
        c.path_offset = lo;
        lo += c.paths;
        c.dt = dt;
        c.precision = precision;
        c.rng = HJM_RNG_PHILOX;
        c.seed = 1234ull;
        c.hazard_rate = 0.0;
        c.recovery = 0.0;
        c.pfe_quantile = 0.0;
        c.device = gi % ndev;
        rc = hjm_plan_create(&plans[gi], mk, sw, &c);
    }
    for (int gi = 0; gi < num_gpus && rc == HJM_OK; ++gi) {
        DeviceGuard guard;
        rc = guard.set(plans[gi]->device);
        if (rc == HJM_OK) rc = plan_launch(plans[gi], false);
    }
    double total[kStatSlots] = {};
    for (int gi = 0; gi < num_gpus && rc == HJM_OK; ++gi) {
        double sums[kStatSlots];
        hjm_results r{};
        r.sums = sums;
        DeviceGuard guard;
        rc = guard.set(plans[gi]->device);
        if (rc == HJM_OK) rc = plan_collect(plans[gi], &r);
        if (rc == HJM_OK)
            for (int t = 0; t < kStatSlots; ++t) total[t] += sums[t];
    }
    for (auto p : plans) hjm_plan_destroy(p);
    if (rc != HJM_OK) return rc;
    const double avg = 1.0 / (double)scenarios;
    for (int t = 0; t < kTenors; ++t) ee[t] = avg * total[t];
    return HJM_OK;
}

int hjm_calculate_exposure_multi_gpu_f64(double* expected_exposure, double notional, double K,
                                         double expiry, double dtau, const double* accrual,
                                         const double* spot_rates, const double* drift,
                                         const double* volatilities, double scale,
                                         int num_gpus, int scenarios, double dt)
{
    (void)scale;
    const hjm_market mk{spot_rates, drift, volatilities, accrual};
    const hjm_swap sw{notional, K, expiry, dtau};
    return multi_gpu_ee(expected_exposure, &mk, &sw, HJM_FP64, num_gpus, scenarios, dt);
}

int hjm_calculate_exposure_multi_gpu_f32(float* expected_exposure, float notional, float K,
                                         float expiry, float dtau, const float* accrual,
                                         const float* spot_rates, const float* drift,
                                         const float* volatilities, float scale,
                                         int num_gpus, int scenarios, float dt)
{
    (void)scale;
    if (!expected_exposure || !accrual || !spot_rates || !drift || !volatilities)
        return HJM_ERR_NULL;
    double acc[kTenors], spot[kTenors], dr[kTenors], vol[kFactors * kTenors], ee[kTenors];
    for (int j = 0; j < kTenors; ++j) {
        acc[j] = accrual[j];
        spot[j] = spot_rates[j];
        dr[j] = drift[j];
    }
    for (int j = 0; j < kFactors * kTenors; ++j) vol[j] = volatilities[j];
    const hjm_market mk{spot, dr, vol, acc};
    const hjm_swap sw{notional, K, expiry, dtau};
    const int rc = multi_gpu_ee(ee, &mk, &sw, HJM_FP32_KAHAN, num_gpus, scenarios, (double)dt);
    if (rc == HJM_OK)
        for (int t = 0; t < kTenors; ++t) expected_exposure[t] = (float)ee[t];
    return rc;
}

}  // extern "C"

// ---- no-allocation variant of hjm_simulate_reference ---------------------------------
// Every intermediate buffer of the run (numeraire pairs num_x, num_y [51][paths] and the
// workgroup partials [blocks][53], all FP64) lives in this static device array, sized for
// HJM_NOALLOC_MAX_PATHS paths: 2 * 51 * 32768 + 128 * 53 doubles = 26.8 MB, held on every
// device the library runs on until the process ends. Nothing in it must start at zero: the
// path kernel writes every numeraire slot the exposure kernel reads (1 .. ndates), and the
// exposure kernel writes every partial row the column sums read.
namespace {

constexpr int64_t kNoallocBlocks = (HJM_NOALLOC_MAX_PATHS + kBlock - 1) / kBlock;
constexpr size_t kScratchDoubles =
    2 * (size_t)kTenors * HJM_NOALLOC_MAX_PATHS + (size_t)kNoallocBlocks * kStatSlots;

static __device__ __align__(256) double scratch[kScratchDoubles];

// The array's address on each device, looked up once per device with that device current
// (every device holds its own copy at its own address). On the host `scratch` names a
// stand-in, never device memory: kernels and copies take these addresses only.
constexpr int kMaxDevices = 64;
std::once_flag g_scratch_once[kMaxDevices];
double* g_scratch_ptr[kMaxDevices];
hipError_t g_scratch_err[kMaxDevices];

int device_scratch(int device, double** out)
{
    if (device < 0 || device >= kMaxDevices) return HJM_ERR_ARG;
    std::call_once(g_scratch_once[device], [device] {
        void* ptr = nullptr;
        g_scratch_err[device] = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        g_scratch_ptr[device] = static_cast<double*>(ptr);
    });
    if (g_scratch_err[device] != hipSuccess || !g_scratch_ptr[device]) return HJM_ERR_DEVICE;
    *out = g_scratch_ptr[device];
    return HJM_OK;
}

}  // namespace

extern "C" int hjm_simulate_reference_noalloc(const hjm_market* market, const hjm_swap* swap,
                                              const hjm_config* cfg, double* d_sums,
                                              void* stream)
{
    int stride, S, D;
    HJM_TRY(check_ref_config(market, swap, cfg, &stride, &S, &D));
    if (!d_sums) return HJM_ERR_NULL;
    if (cfg->paths > HJM_NOALLOC_MAX_PATHS) return HJM_ERR_LIMIT;
    if (cfg->rng != HJM_RNG_PHILOX || cfg->path_kernel == HJM_KERNEL_MATRIX)
        return HJM_ERR_UNSUPPORTED;
    RefStepConsts<double> kd{};
    RefStepConsts<float> kf{};
    RefGrid g{};
    ExposureParams e{};
    fill_ref_params(market, swap, cfg, stride, S, D, kd, kf, g, e);

    DeviceGuard guard;
    HJM_TRY(guard.set(cfg->device));
    double* base = nullptr;
    HJM_TRY(device_scratch(cfg->device, &base));
    const int64_t P = cfg->paths, nblocks = (P + kBlock - 1) / kBlock;
    double* nx = base;
    double* ny = base + (size_t)kTenors * P;
    double* partials = base + 2 * (size_t)kTenors * P;
    hipStream_t st = static_cast<hipStream_t>(stream);
    HJM_HIP(cfg->precision == HJM_FP64 ? launch_ref_paths_f64(kd, g, nullptr, nx, ny, st)
                                       : launch_ref_paths_f32k(kf, g, nullptr, nx, ny, st),
            HJM_ERR_LAUNCH);
    HJM_HIP(launch_ref_exposure(e, nx, ny, nullptr, partials, st), HJM_ERR_LAUNCH);
    HJM_HIP(launch_column_sum(partials, nblocks, kStatSlots, d_sums, st), HJM_ERR_LAUNCH);
    return HJM_OK;
}
