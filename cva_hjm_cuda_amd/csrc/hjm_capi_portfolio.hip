// hjm_capi_portfolio.hip -- C ABI of the portfolio semantics: coefficient tables,
// persistent-grid tuning, hipGraph capture and replay, autotuner.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "hjm_host.h"
#include "hjm_kernels.h"
#include "hjm_mfma_plan.h"
#include "hjm_select.h"

using namespace hjm;
using namespace hjm_host;

struct hjm_pplan_s {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[5] = {};
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hjm_tuning tuning{256, 0, 0, 0, 0, 0, 0, 0, 0, 0};  // block size 256, the rest defaults
    MfmaShape path_shape{}, value_shape{};             // matrix-core kernels, from tuning
    int cus = 0;
    int precision = HJM_FP64;
    bool do_pfe = false;
    int64_t paths = 0;
    int nsets = 0, ndates = 0;
    int64_t pfe_rank = 1;
    int grid_blocks = 0;
    bool use_mfma = false;    // FP32 on the matrix cores (path kernel and valuation)
    bool use_mfma64 = false;  // FP64 path kernel on the matrix cores (HJM_KERNEL_MATRIX)
    MfmaGrid mg{};
    uint16_t* d_wfrag = nullptr;  // bf16 parts of the block map
    double* d_wfrag64 = nullptr;  // the block map's FP64 fragments
    RefStepConsts<double> kd{};
    RefStepConsts<float> kf{};
    PfGrid grid{};
    PfValueParams vp{};
    std::vector<double> w;  // [nsets][ndates] CVA weights
    double* d_normals = nullptr;
    void* d_curves = nullptr;
    double* d_logb = nullptr;
    void* d_ex = nullptr;
    double* d_partials = nullptr;
    double* d_sums = nullptr;
    double* d_pfe = nullptr;
    char* d_selws = nullptr;  // bracket-selection workspace (PFE)
    uint32_t* d_loparts = nullptr;  // counts below the bracket per row and statistics workgroup
    void* d_coef = nullptr;
    void* d_a = nullptr;
    int* d_kmax = nullptr;
    double* d_w = nullptr;
    uint16_t* d_vfrag = nullptr;  // FP32: bf16 parts of -gamma (matrix-core valuation)
    float* d_apad = nullptr;
    int* d_tilek = nullptr;
};

namespace {

void pplan_release(hjm_portfolio_plan p)
{
    if (!p) return;
    int prev = -1;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(p->device);
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    dev_free(p->d_normals);
    dev_free(p->d_curves);
    dev_free(p->d_logb);
    dev_free(p->d_ex);
    dev_free(p->d_partials);
    dev_free(p->d_sums);
    dev_free(p->d_pfe);
    dev_free(p->d_selws);
    dev_free(p->d_loparts);
    dev_free(p->d_coef);
    dev_free(p->d_vfrag);
    dev_free(p->d_apad);
    dev_free(p->d_tilek);
    dev_free(p->d_a);
    dev_free(p->d_kmax);
    dev_free(p->d_w);
    dev_free(p->d_wfrag);
    dev_free(p->d_wfrag64);
    for (auto& e : p->ev)
        if (e) (void)hipEventDestroy(e);
    if (p->stream) (void)hipStreamDestroy(p->stream);
    if (prev >= 0) (void)hipSetDevice(prev);
    delete p;
}

int check_portfolio(const hjm_portfolio* pf, int* ndates)
{
    if (!pf || !pf->periods || !pf->sign_notional || !pf->fixed_rate || !pf->netting_set ||
        !pf->hazard_rate)
        return HJM_ERR_NULL;
    if (pf->nswaps <= 0 || pf->nswaps > HJM_MAX_SWAPS || pf->nsets <= 0 || pf->nsets > HJM_MAX_SETS)
        return HJM_ERR_ARG;
    if (!(pf->dtau > 0.0) || !(pf->horizon > 0.0) || !(pf->recovery >= 0.0 && pf->recovery <= 1.0))
        return HJM_ERR_ARG;
    const double r = pf->horizon / pf->dtau;
    const int E = (int)llround(r);
    if (E < 1 || fabs(r - E) > 1e-9 * E || E > kMaxPeriods) return HJM_ERR_GRID;
    for (int s = 0; s < pf->nswaps; ++s) {
        if (pf->periods[s] < 1 || pf->periods[s] > E) return HJM_ERR_ARG;
        if (pf->netting_set[s] < 0 || pf->netting_set[s] >= pf->nsets) return HJM_ERR_ARG;
        if (!isfinite(pf->sign_notional[s]) || !isfinite(pf->fixed_rate[s])) return HJM_ERR_ARG;
    }
    for (int g = 0; g < pf->nsets; ++g)
        if (!(pf->hazard_rate[g] >= 0.0)) return HJM_ERR_ARG;
    *ndates = E + 1;
    return HJM_OK;
}

// V_g(t_e) = a_ge - sum_k gamma_gek P_k with, for every live swap (k_s = n_s - e > 0):
// a += sN, gamma_{k_s} += sN (the -sN P_{n} term), gamma_j += sN K dtau for j <= k_s.
void build_tables(const hjm_portfolio* pf, int ndates, std::vector<double>& coef,
                  std::vector<double>& a, std::vector<int>& kmax)
{
    const int G = pf->nsets;
    coef.assign((size_t)ndates * G * kMaxPeriods, 0.0);
    a.assign((size_t)ndates * G, 0.0);
    kmax.assign((size_t)ndates * G, 0);
    for (int e = 0; e < ndates; ++e) {
        for (int s = 0; s < pf->nswaps; ++s) {
            const int k = pf->periods[s] - e;
            if (k <= 0) continue;
            const int g = pf->netting_set[s];
            const size_t row = (size_t)e * G + g;
            const double sn = pf->sign_notional[s];
            a[row] += sn;
            double* gam = &coef[row * kMaxPeriods];
            gam[k - 1] += sn;
            const double fx = sn * pf->fixed_rate[s] * pf->dtau;
            for (int j = 1; j <= k; ++j) gam[j - 1] += fx;
            kmax[row] = std::max(kmax[row], k);
        }
    }
}

int check_tuning(const hjm_tuning& t)
{
    if (t.block_size != 64 && t.block_size != 128 && t.block_size != 256) return HJM_ERR_ARG;
    if (t.blocks_per_cu < 0 || t.blocks_per_cu > 64) return HJM_ERR_ARG;
    if (t.use_graph != 0 && t.use_graph != 1) return HJM_ERR_ARG;
    if (t.fuse_statistics != 0 && t.fuse_statistics != 1) return HJM_ERR_ARG;
    if (t.pfe_stage_slots < 0 || t.pfe_stage_slots > 32) return HJM_ERR_ARG;
    if (!pf_path_shape_ok(t.path_tiles ? t.path_tiles : 2, t.path_waves ? t.path_waves : 8))
        return HJM_ERR_ARG;
    if (!pf_value_shape_ok(t.value_tiles ? t.value_tiles : 4, t.value_waves ? t.value_waves : 4))
        return HJM_ERR_ARG;
    if (t.value_stage < 0 || t.value_stage > 8) return HJM_ERR_ARG;
    return HJM_OK;
}

// the kernels' launch shapes for a (checked) tuning
void pplan_apply_tuning(hjm_portfolio_plan p, const hjm_tuning& t)
{
    p->tuning = t;
    const int64_t tiles = (p->paths + t.block_size - 1) / t.block_size;
    p->grid_blocks = (int)std::min<int64_t>(
        tiles, t.blocks_per_cu > 0 ? (int64_t)t.blocks_per_cu * p->cus : tiles);
    p->path_shape = MfmaShape{t.path_tiles, t.path_waves, 0,
                              t.blocks_per_cu > 0 ? (int64_t)t.blocks_per_cu * p->cus : 0};
    p->value_shape = MfmaShape{t.value_tiles, t.value_waves, t.value_stage, 0};
}

int pplan_init(hjm_portfolio_plan p, const hjm_market* market, const hjm_portfolio* pf,
               const hjm_config* cfg, const hjm_tuning* tuning)
{
    HJM_TRY(check_market(market));
    if (!cfg) return HJM_ERR_NULL;
    int nd = 0;
    HJM_TRY(check_portfolio(pf, &nd));
    if (cfg->paths <= 0 || cfg->path_offset < 0 || cfg->path_offset + cfg->paths > (1ll << 32))
        return HJM_ERR_ARG;
    if (cfg->precision != HJM_FP64 && cfg->precision != HJM_FP32_KAHAN) return HJM_ERR_ARG;
    if (cfg->rng != HJM_RNG_PHILOX && cfg->rng != HJM_RNG_HOST) return HJM_ERR_ARG;
    if (cfg->rng == HJM_RNG_HOST && !cfg->host_normals) return HJM_ERR_NULL;
    if (cfg->pfe_quantile > 1.0) return HJM_ERR_ARG;
    if (tuning) {
        HJM_TRY(check_tuning(*tuning));
        p->tuning = *tuning;
    }
    if (!(cfg->dt > 0.0)) return HJM_ERR_ARG;
    const double rs = pf->dtau / cfg->dt;
    const int stride = (int)llround(rs);
    if (stride <= 0 || fabs(rs - stride) > 1e-6 * stride) return HJM_ERR_GRID;
    const int S = stride * (nd - 1);
    if (cfg->path_kernel < HJM_KERNEL_AUTO || cfg->path_kernel > HJM_KERNEL_MATRIX) return HJM_ERR_ARG;
    if (cfg->adjoint_normals != HJM_ADJ_NORMALS_AUTO && cfg->adjoint_normals != HJM_ADJ_NORMALS_REGENERATE)
        return HJM_ERR_ARG;
    const bool mfma_ok = cfg->rng == HJM_RNG_PHILOX && stride == kMfmaStride;
    if (cfg->path_kernel == HJM_KERNEL_MATRIX && !mfma_ok) return HJM_ERR_UNSUPPORTED;
    // AUTO: the matrix cores for FP32; FP64 keeps the vector path kernel unless MATRIX is
    // asked for (its valuation and statistics stay on the vector ALUs either way)
    p->use_mfma = mfma_ok && cfg->precision == HJM_FP32_KAHAN && cfg->path_kernel != HJM_KERNEL_VECTOR;
    p->use_mfma64 = mfma_ok && cfg->precision == HJM_FP64 && cfg->path_kernel == HJM_KERNEL_MATRIX;

    p->device = cfg->device;
    p->precision = cfg->precision;
    p->paths = cfg->paths;
    p->nsets = pf->nsets;
    p->ndates = nd;
    p->do_pfe = cfg->pfe_quantile > 0.0;
    p->pfe_rank = quantile_rank(cfg->paths, cfg->pfe_quantile);
    fill_consts(p->kd, market, cfg->dt, pf->dtau);
    fill_consts(p->kf, market, cfg->dt, pf->dtau);
    PfGrid& g = p->grid;
    g.stride = stride;
    g.nsteps = S;
    g.ndates_m1 = nd - 1;
    g.paths = cfg->paths;
    g.path0 = cfg->path_offset;
    g.seed_lo = (uint32_t)cfg->seed;
    g.seed_hi = (uint32_t)(cfg->seed >> 32);
    g.dt = cfg->dt;
    p->vp.paths = cfg->paths;
    p->vp.nsets = pf->nsets;
    p->vp.ndates = nd;
    p->vp.dtau = pf->dtau;
    p->w.assign((size_t)pf->nsets * nd, 0.0);
    for (int s = 0; s < pf->nsets; ++s)
        cva_weights(nd, pf->dtau, pf->hazard_rate[s], pf->recovery, &p->w[(size_t)s * nd]);
    std::vector<double> coef, a;
    std::vector<int> kmax;
    build_tables(pf, nd, coef, a, kmax);

    DeviceGuard guard;
    HJM_TRY(guard.set(cfg->device));
    HJM_HIP(hipDeviceGetAttribute(&p->cus, hipDeviceAttributeMultiprocessorCount, cfg->device),
            HJM_ERR_DEVICE);
    pplan_apply_tuning(p, p->tuning);
    HJM_HIP(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking), HJM_ERR_DEVICE);
    for (auto& ev : p->ev) HJM_HIP(hipEventCreate(&ev), HJM_ERR_DEVICE);
    const size_t rsz = (p->precision == HJM_FP64) ? sizeof(double) : sizeof(float);
    const size_t P = (size_t)cfg->paths;
    // workgroups of the statistics: 256 paths (exposure pass) or the fused pass's blocks
    const size_t nb = std::max<size_t>((P + kBlock - 1) / kBlock, (size_t)pf_fused_blocks(cfg->paths));
    char* raw = nullptr;
    HJM_TRY(dev_alloc(&raw, (size_t)nd * kTenors * P * rsz));
    p->d_curves = raw;
    HJM_TRY(dev_alloc(&p->d_logb, (size_t)nd * P));
    raw = nullptr;
    HJM_TRY(dev_alloc(&raw, (size_t)pf->nsets * nd * P * rsz));
    p->d_ex = raw;
    HJM_TRY(dev_alloc(&p->d_partials, (size_t)pf->nsets * nb * (nd + 2)));
    HJM_TRY(dev_alloc(&p->d_sums, (size_t)pf->nsets * (nd + 2)));
    HJM_TRY(dev_alloc(&p->d_pfe, (size_t)pf->nsets * nd));
    if (p->do_pfe) {
        const size_t wsb = bracket_workspace_bytes(pf->nsets * nd, bracket_plan(cfg->paths, p->pfe_rank, (int)rsz),
                                                   (int)rsz);
        if (wsb) {
            HJM_TRY(dev_alloc(&p->d_selws, wsb));
            if (rsz == sizeof(float))
                HJM_TRY(dev_alloc(&p->d_loparts, (size_t)pf->nsets * nd * pf_stats_blocks(cfg->paths)));
        }
    }
    HJM_TRY(dev_alloc(&p->d_kmax, kmax.size()));
    HJM_TRY(dev_alloc(&p->d_w, p->w.size()));
    HJM_HIP(hipMemcpy(p->d_kmax, kmax.data(), kmax.size() * sizeof(int), hipMemcpyHostToDevice),
            HJM_ERR_DEVICE);
    p->vp.kmax = p->d_kmax;
    // rows of matured sets are never written (their exposures are 0) and never read by the
    // statistics pass; they hold zeros for the PFE selection and the per-path copy
    HJM_HIP(hipMemset(p->d_ex, 0, (size_t)pf->nsets * nd * P * rsz), HJM_ERR_DEVICE);
    HJM_HIP(hipMemcpy(p->d_w, p->w.data(), p->w.size() * sizeof(double), hipMemcpyHostToDevice),
            HJM_ERR_DEVICE);
    if (p->precision == HJM_FP64) {
        double *c = nullptr, *aa = nullptr;
        HJM_TRY(dev_alloc(&c, coef.size()));
        p->d_coef = c;
        HJM_TRY(dev_alloc(&aa, a.size()));
        p->d_a = aa;
        HJM_HIP(hipMemcpy(c, coef.data(), coef.size() * sizeof(double), hipMemcpyHostToDevice),
                HJM_ERR_DEVICE);
        HJM_HIP(hipMemcpy(aa, a.data(), a.size() * sizeof(double), hipMemcpyHostToDevice),
                HJM_ERR_DEVICE);
    } else {
        std::vector<float> cf(coef.begin(), coef.end()), af(a.begin(), a.end());
        float *c = nullptr, *aa = nullptr;
        HJM_TRY(dev_alloc(&c, cf.size()));
        p->d_coef = c;
        HJM_TRY(dev_alloc(&aa, af.size()));
        p->d_a = aa;
        HJM_HIP(hipMemcpy(c, cf.data(), cf.size() * sizeof(float), hipMemcpyHostToDevice),
                HJM_ERR_DEVICE);
        HJM_HIP(hipMemcpy(aa, af.data(), af.size() * sizeof(float), hipMemcpyHostToDevice),
                HJM_ERR_DEVICE);
        std::vector<uint16_t> vf;
        std::vector<float> ap;
        std::vector<int> tk;
        build_value_fragments(coef, a, kmax, nd, pf->nsets, kMaxPeriods, vf, ap, tk);
        HJM_TRY(dev_alloc(&p->d_vfrag, vf.size()));
        HJM_TRY(dev_alloc(&p->d_apad, ap.size()));
        HJM_TRY(dev_alloc(&p->d_tilek, tk.size()));
        HJM_HIP(hipMemcpy(p->d_vfrag, vf.data(), vf.size() * sizeof(uint16_t), hipMemcpyHostToDevice),
                HJM_ERR_DEVICE);
        HJM_HIP(hipMemcpy(p->d_apad, ap.data(), ap.size() * sizeof(float), hipMemcpyHostToDevice),
                HJM_ERR_DEVICE);
        HJM_HIP(hipMemcpy(p->d_tilek, tk.data(), tk.size() * sizeof(int), hipMemcpyHostToDevice),
                HJM_ERR_DEVICE);
    }
    if (cfg->rng == HJM_RNG_HOST) {
        const size_t n = P * (size_t)S * 3;
        HJM_TRY(dev_alloc(&p->d_normals, n));
        HJM_HIP(hipMemcpy(p->d_normals, cfg->host_normals, n * sizeof(double), hipMemcpyHostToDevice),
                HJM_ERR_DEVICE);
    }
    if (p->use_mfma) {
        HJM_HIP(mfma_kernels_prepare(), HJM_ERR_DEVICE);
        std::vector<uint16_t> wf;
        build_mfma_block_map(market->drifts, market->volatilities, cfg->dt, pf->dtau, wf,
                             MfmaMap::kPortfolio);
        HJM_TRY(dev_alloc(&p->d_wfrag, wf.size()));
        HJM_HIP(hipMemcpy(p->d_wfrag, wf.data(), wf.size() * sizeof(uint16_t), hipMemcpyHostToDevice),
                HJM_ERR_DEVICE);
    }
    if (p->use_mfma64) {
        HJM_HIP(mfma64_kernels_prepare(), HJM_ERR_DEVICE);
        std::vector<double> wf;
        build_mfma64_block_map(market->drifts, market->volatilities, cfg->dt, pf->dtau,
                               kMfmaPfMTiles, wf, MfmaMap::kPortfolio);
        HJM_TRY(dev_alloc(&p->d_wfrag64, wf.size()));
        HJM_HIP(hipMemcpy(p->d_wfrag64, wf.data(), wf.size() * sizeof(double), hipMemcpyHostToDevice),
                HJM_ERR_DEVICE);
    }
    if (p->use_mfma || p->use_mfma64) {
        MfmaGrid& mg = p->mg;
        mg.ndates = nd;
        mg.paths = cfg->paths;
        mg.path0 = cfg->path_offset;
        mg.seed_lo = g.seed_lo;
        mg.seed_hi = g.seed_hi;
        mg.dt = cfg->dt;
        for (int j = 0; j < kTenors; ++j) {
            mg.spot[j] = (float)market->spot_rates[j];
            mg.spot64[j] = market->spot_rates[j];
        }
    }
    return HJM_OK;
}

// FP32 plans value on the matrix cores. With tuning.fuse_statistics, unless the caller
// wants the per-path exposures (or the PFE needs whole rows: fewer than 4 x 16384 paths),
// valuation and statistics run fused, without writing the exposures: the PFE's bracket
// comes from the first 16384 paths valued on their own, the fused pass counts and
// collects every row's candidates, and only rows the counting leaves unresolved are
// valued again, whole, for the selection.
bool pplan_fused(hjm_portfolio_plan p, bool keep_exposures)
{
    if (p->precision == HJM_FP64 || keep_exposures || !p->tuning.fuse_statistics) return false;
    return !p->do_pfe || (p->d_selws && bracket_plan(p->paths, p->pfe_rank, (int)sizeof(float)).use);
}

// Enqueue the passes on the plan's stream, with timing events between them unless the
// sequence is being captured into a graph (a captured replay is timed as a whole).
// Timings: [0] paths, [1] valuation (+ statistics when fused), [2] statistics (the column
// sums when fused), [3] PFE selection.
int pplan_enqueue(hjm_portfolio_plan p, bool events = true, float* zs = nullptr,
                  bool keep_exposures = false)
{
    hipStream_t st = p->stream;
    const bool f64 = p->precision == HJM_FP64;
    const int rows = p->nsets * p->ndates;
    auto mark = [&](int i) { return events ? hipEventRecord(p->ev[i], st) : hipSuccess; };
    HJM_HIP(mark(0), HJM_ERR_DEVICE);
    hipError_t err;
    if (p->use_mfma)
        err = launch_pf_paths_mfma(p->mg, p->d_wfrag, (float*)p->d_curves, p->d_logb, zs,
                                   p->path_shape, st);
    else if (p->use_mfma64)
        err = launch_pf_paths_mfma64(p->mg, p->d_wfrag64, (double*)p->d_curves, p->d_logb, zs, st);
    else if (f64)
        err = launch_pf_paths_f64(p->kd, p->grid, p->d_normals, (double*)p->d_curves, p->d_logb,
                                  p->tuning.block_size, p->grid_blocks, st);
    else
        err = launch_pf_paths_f32(p->kf, p->grid, p->d_normals, (float*)p->d_curves, p->d_logb,
                                  p->tuning.block_size, p->grid_blocks, st);
    HJM_HIP(err, HJM_ERR_LAUNCH);
    HJM_HIP(mark(1), HJM_ERR_DEVICE);
    if (pplan_fused(p, keep_exposures)) {
        const float* curves = (const float*)p->d_curves;
        float* ex = (float*)p->d_ex;
        SelRowState* state = nullptr;
        float* cand = nullptr;
        BracketPlan bp{};
        if (p->do_pfe) {
            bp = bracket_plan(p->paths, p->pfe_rank, (int)sizeof(float));
            bracket_ws_split<float>(p->d_selws, rows, &state, &cand);
            // the sample: exposures of the first bp.m paths, then the rows' brackets
            HJM_HIP(launch_pf_value_mfma(p->vp, curves, p->d_logb, p->d_vfrag, p->d_apad, p->d_tilek,
                                         ex, p->value_shape, st, bp.m),
                    HJM_ERR_LAUNCH);
            HJM_HIP(launch_bracket_sample_f32(ex, rows, p->paths, p->pfe_rank, p->d_selws, st),
                    HJM_ERR_LAUNCH);
        }
        HJM_HIP(launch_pf_value_stats_mfma(p->vp, curves, p->d_logb, p->d_vfrag, p->d_apad,
                                           p->d_tilek, p->d_w, p->d_partials, state, cand, bp.cap,
                                           p->tuning.pfe_stage_slots, st),
                HJM_ERR_LAUNCH);
        HJM_HIP(mark(2), HJM_ERR_DEVICE);
        HJM_HIP(launch_grouped_column_sum(p->d_partials, p->nsets, pf_fused_blocks(p->paths),
                                          p->ndates + 2, p->d_sums, st),
                HJM_ERR_LAUNCH);
        HJM_HIP(mark(3), HJM_ERR_DEVICE);
        if (p->do_pfe) {
            // whole rows only where the counting did not resolve the rank
            HJM_HIP(launch_pf_value_mfma(p->vp, curves, p->d_logb, p->d_vfrag, p->d_apad, p->d_tilek,
                                         ex, p->value_shape, st, -1, state, p->pfe_rank, bp.cap),
                    HJM_ERR_LAUNCH);
            HJM_HIP(launch_bracket_finish_f32(ex, rows, p->paths, p->pfe_rank, p->d_selws, p->d_pfe, st),
                    HJM_ERR_LAUNCH);
        }
        HJM_HIP(mark(4), HJM_ERR_DEVICE);
        return HJM_OK;
    }
    err = f64 ? launch_pf_value_f64(p->vp, (const double*)p->d_curves, p->d_logb,
                                    (const double*)p->d_coef, (const double*)p->d_a, p->d_kmax,
                                    (double*)p->d_ex, st)
              : launch_pf_value_mfma(p->vp, (const float*)p->d_curves, p->d_logb, p->d_vfrag,
                                     p->d_apad, p->d_tilek, (float*)p->d_ex, p->value_shape, st);
    HJM_HIP(err, HJM_ERR_LAUNCH);
    HJM_HIP(mark(2), HJM_ERR_DEVICE);
    // With a bracket-selection workspace the statistics pass also does the PFE's counting
    // pass: the bracket sample runs before it and the selection's last launch after it.
    void* selws = (p->do_pfe && p->d_selws &&
                   pf_stats_stage(p->paths, p->pfe_rank, p->ndates, f64 ? 8 : 4) > 0)
                      ? p->d_selws
                      : nullptr;
    if (selws) {
        err = f64 ? launch_bracket_sample((const double*)p->d_ex, rows, p->paths, p->pfe_rank, selws, st)
                  : launch_bracket_sample_f32((const float*)p->d_ex, rows, p->paths, p->pfe_rank, selws,
                                              st);
        HJM_HIP(err, HJM_ERR_LAUNCH);
    }
    err = f64 ? launch_pf_stats_f64(p->vp, (const double*)p->d_ex, p->d_w, p->d_partials,
                                    p->d_sums, selws, p->pfe_rank, nullptr, st)
              : launch_pf_stats_f32(p->vp, (const float*)p->d_ex, p->d_w, p->d_partials,
                                    p->d_sums, selws, p->pfe_rank, p->d_loparts, st);
    HJM_HIP(err, HJM_ERR_LAUNCH);
    HJM_HIP(mark(3), HJM_ERR_DEVICE);
    if (p->do_pfe) {
        if (selws)
            err = f64 ? launch_bracket_finish((const double*)p->d_ex, rows, p->paths, p->pfe_rank,
                                              selws, p->d_pfe, st)
                      : launch_bracket_finish_f32((const float*)p->d_ex, rows, p->paths, p->pfe_rank,
                                                  selws, p->d_pfe, st);
        else
            err = f64 ? launch_radix_select((const double*)p->d_ex, rows, p->paths, p->pfe_rank,
                                            p->d_selws, p->d_pfe, st)
                      : launch_radix_select_f32((const float*)p->d_ex, rows, p->paths, p->pfe_rank,
                                                p->d_selws, p->d_pfe, st);
        HJM_HIP(err, HJM_ERR_LAUNCH);
    }
    HJM_HIP(mark(4), HJM_ERR_DEVICE);
    return HJM_OK;
}

int pplan_run(hjm_portfolio_plan p)
{
    if (!p->tuning.use_graph) return pplan_enqueue(p);
    if (!p->exec) {
        HJM_HIP(hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal), HJM_ERR_DEVICE);
        const int rc = pplan_enqueue(p, false);
        hipGraph_t graph = nullptr;
        const hipError_t end = hipStreamEndCapture(p->stream, &graph);
        if (rc != HJM_OK) {
            if (graph) (void)hipGraphDestroy(graph);
            return rc;
        }
        HJM_HIP(end, HJM_ERR_DEVICE);
        p->graph = graph;
        HJM_HIP(hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0), HJM_ERR_DEVICE);
    }
    HJM_HIP(hipEventRecord(p->ev[0], p->stream), HJM_ERR_DEVICE);
    HJM_HIP(hipGraphLaunch(p->exec, p->stream), HJM_ERR_LAUNCH);
    HJM_HIP(hipEventRecord(p->ev[4], p->stream), HJM_ERR_DEVICE);
    return HJM_OK;
}

int pplan_collect(hjm_portfolio_plan p, hjm_portfolio_results* r)
{
    HJM_HIP(hipStreamSynchronize(p->stream), HJM_ERR_LAUNCH);
    const int G = p->nsets, nd = p->ndates, nc = nd + 2;
    std::vector<double> sums((size_t)G * nc);
    HJM_HIP(hipMemcpy(sums.data(), p->d_sums, sums.size() * sizeof(double), hipMemcpyDeviceToHost),
            HJM_ERR_DEVICE);
    const double P = (double)p->paths, inv = 1.0 / P;
    for (int g = 0; g < G; ++g) {
        const double* sg = &sums[(size_t)g * nc];
        double cva = 0.0;
        for (int e = 0; e < nd; ++e) {
            const double ee = inv * sg[e];
            if (r->expected_exposure) r->expected_exposure[(size_t)g * nd + e] = ee;
            cva += p->w[(size_t)g * nd + e] * ee;
        }
        if (r->cva) r->cva[g] = cva;
        if (r->cva_stderr) {
            const double s1 = sg[nd], s2 = sg[nd + 1];
            const double var = (p->paths > 1) ? (s2 - s1 * s1 / P) / (P - 1.0) : 0.0;
            r->cva_stderr[g] = sqrt((var > 0.0 ? var : 0.0) / P);
        }
    }
    // device rows are [date][set]; the results are [set][date]
    if (r->pfe) {
        if (p->do_pfe) {
            std::vector<double> pf((size_t)G * nd);
            HJM_HIP(hipMemcpy(pf.data(), p->d_pfe, sizeof(double) * G * nd, hipMemcpyDeviceToHost),
                    HJM_ERR_DEVICE);
            for (int e = 0; e < nd; ++e)
                for (int g = 0; g < G; ++g) r->pfe[(size_t)g * nd + e] = pf[(size_t)e * G + g];
        } else {
            for (int i = 0; i < G * nd; ++i) r->pfe[i] = 0.0;
        }
    }
    if (r->exposures) {
        const size_t cells = (size_t)G * nd * p->paths;
        std::vector<double> tmp(cells);
        if (p->precision == HJM_FP64) {
            HJM_HIP(hipMemcpy(tmp.data(), p->d_ex, cells * sizeof(double), hipMemcpyDeviceToHost),
                    HJM_ERR_DEVICE);
        } else {
            std::vector<float> tf(cells);
            HJM_HIP(hipMemcpy(tf.data(), p->d_ex, cells * sizeof(float), hipMemcpyDeviceToHost),
                    HJM_ERR_DEVICE);
            for (size_t i = 0; i < cells; ++i) tmp[i] = tf[i];
        }
        for (int e = 0; e < nd; ++e)
            for (int g = 0; g < G; ++g) {
                const double* src = &tmp[((size_t)e * G + g) * p->paths];
                for (int64_t i = 0; i < p->paths; ++i)
                    r->exposures[((size_t)i * G + g) * nd + e] = src[i];
            }
    }
    if (r->timings_ms) {
        float ms = 0.f;
        for (int i = 0; i < 4; ++i) {
            if (p->tuning.use_graph) {
                r->timings_ms[i] = NAN;  // one graph replay: only its total is timed
                continue;
            }
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

int hjm_portfolio_plan_create(hjm_portfolio_plan* out, const hjm_market* market,
                              const hjm_portfolio* portfolio, const hjm_config* config,
                              const hjm_tuning* tuning)
{
    if (!out) return HJM_ERR_NULL;
    *out = nullptr;
    hjm_portfolio_plan p = new (std::nothrow) hjm_pplan_s();
    if (!p) return HJM_ERR_ALLOC;
    const int rc = pplan_init(p, market, portfolio, config, tuning);
    if (rc != HJM_OK) {
        pplan_release(p);
        return rc;
    }
    *out = p;
    return HJM_OK;
}

int hjm_portfolio_plan_execute(hjm_portfolio_plan plan, hjm_portfolio_results* results)
{
    if (!plan || !results) return HJM_ERR_NULL;
    DeviceGuard guard;
    HJM_TRY(guard.set(plan->device));
    // per-path exposures asked for: an eager run that writes them (the graph, if any, is
    // the fused form)
    if (results->exposures) HJM_TRY(pplan_enqueue(plan, true, nullptr, true));
    else HJM_TRY(pplan_run(plan));
    return pplan_collect(plan, results);
}

void hjm_portfolio_plan_destroy(hjm_portfolio_plan plan) { pplan_release(plan); }

int hjm_simulate_portfolio(const hjm_market* market, const hjm_portfolio* portfolio,
                           const hjm_config* config, const hjm_tuning* tuning,
                           hjm_portfolio_results* results)
{
    if (!results) return HJM_ERR_NULL;
    hjm_portfolio_plan p = nullptr;
    HJM_TRY(hjm_portfolio_plan_create(&p, market, portfolio, config, tuning));
    const int rc = hjm_portfolio_plan_execute(p, results);
    hjm_portfolio_plan_destroy(p);
    return rc;
}

int hjm_portfolio_cva_greeks(const hjm_market* market, const hjm_portfolio* portfolio,
                             const hjm_config* config, const hjm_tuning* tuning, hjm_greeks* gr)
{
    if (!gr || !config) return HJM_ERR_NULL;
    hjm_config cfg = *config;
    cfg.pfe_quantile = 0.0;
    hjm_portfolio_plan p = nullptr;
    HJM_TRY(hjm_portfolio_plan_create(&p, market, portfolio, &cfg, tuning));
    constexpr int kOut = 5 * kTenors;  // g0, g1, g2, sum of lambda_e, d spot
    struct Cleanup {
        hjm_portfolio_plan p;
        float *gbar = nullptr, *beta = nullptr, *zs = nullptr;
        double *part = nullptr, *sums = nullptr, *bp = nullptr;
        uint16_t* tfrag = nullptr;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        ~Cleanup()
        {
            dev_free(zs);
            dev_free(gbar);
            dev_free(beta);
            dev_free(part);
            dev_free(sums);
            dev_free(bp);
            dev_free(tfrag);
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
            hjm_portfolio_plan_destroy(p);
        }
    } c{p};
    DeviceGuard guard;
    HJM_TRY(guard.set(p->device));
    const int64_t P = p->paths;
    const int nd = p->ndates, E = nd - 1;
    const int64_t nb = pf_adjoint_blocks(P);
    HJM_TRY(dev_alloc(&c.gbar, (size_t)nd * kTenors * P));
    HJM_TRY(dev_alloc(&c.beta, (size_t)nd * P));
    HJM_TRY(dev_alloc(&c.part, (size_t)nb * kOut));
    HJM_TRY(dev_alloc(&c.sums, (size_t)kOut + kMaxDates));  // then the per-date sums of beta
    const bool f64 = p->precision == HJM_FP64;
    // rows of the seeds' beta partials: blocks of 256 paths (FP64) or the matrix-core
    // kernel's workgroups (FP32)
    const int64_t tiles = f64 ? (P + kBlock - 1) / kBlock : pf_adjoint_seed_mfma_blocks(P);
    HJM_TRY(dev_alloc(&c.bp, (size_t)tiles * nd));
    if (!f64) {
        // -Gamma^T in the order the seed's second product reads it
        std::vector<double> coef, a;
        std::vector<int> kmax;
        build_tables(portfolio, nd, coef, a, kmax);
        std::vector<uint16_t> tf;
        build_adjoint_fragments(coef, nd, p->nsets, kMaxPeriods, tf);
        HJM_TRY(dev_alloc(&c.tfrag, tf.size()));
        HJM_HIP(hipMemcpy(c.tfrag, tf.data(), tf.size() * sizeof(uint16_t), hipMemcpyHostToDevice),
                HJM_ERR_DEVICE);
    }
    HJM_HIP(hipEventCreate(&c.e0), HJM_ERR_DEVICE);
    HJM_HIP(hipEventCreate(&c.e1), HJM_ERR_DEVICE);
    // the matrix-core forward keeps its normals for the reverse sweep; without the memory
    // the sweep regenerates them
    if ((p->use_mfma || p->use_mfma64) && config->adjoint_normals == HJM_ADJ_NORMALS_AUTO &&
        dev_alloc(&c.zs, (size_t)p->grid.nsteps * 3 * (size_t)P) != HJM_OK)
        c.zs = nullptr;
    hipStream_t st = p->stream;
    HJM_TRY(pplan_enqueue(p, true, c.zs));
    HJM_HIP(hipEventRecord(c.e0, st), HJM_ERR_DEVICE);
    hipError_t err = f64 ? launch_pf_adjoint_seed_f64(p->vp, (const double*)p->d_curves, p->d_logb,
                                                      (const double*)p->d_coef, (const double*)p->d_a,
                                                      p->d_kmax, p->d_w, c.gbar, c.beta, c.bp, st)
                         : launch_pf_adjoint_seed_mfma(p->vp, (const float*)p->d_curves, p->d_logb,
                                                       p->d_vfrag, c.tfrag, p->d_apad, p->d_tilek,
                                                       p->d_w, c.gbar, c.beta, c.bp, st);
    HJM_HIP(err, HJM_ERR_LAUNCH);
    const double cc = p->grid.dt / portfolio->dtau;
    // beta summed over the paths at every date (the drift gradient's bank-account terms)
    HJM_HIP(launch_column_sum(c.bp, tiles, nd, c.sums + kOut, st), HJM_ERR_LAUNCH);
    HJM_HIP(launch_pf_adjoint(p->grid, cc, p->d_normals, c.zs, c.gbar, c.beta, c.part, st),
            HJM_ERR_LAUNCH);
    HJM_HIP(launch_column_sum(c.part, nb, kOut, c.sums, st), HJM_ERR_LAUNCH);
    HJM_HIP(hipEventRecord(c.e1, st), HJM_ERR_DEVICE);
    std::vector<double> cvas(p->nsets);
    hjm_portfolio_results r{};
    r.cva = cvas.data();
    HJM_TRY(pplan_collect(p, &r));
    std::vector<double> sums((size_t)kOut + kMaxDates, 0.0);
    HJM_HIP(hipMemcpy(sums.data(), c.sums, sizeof(double) * (kOut + nd), hipMemcpyDeviceToHost),
            HJM_ERR_DEVICE);
    double total = 0.0;
    for (double v : cvas) total += v;
    if (gr->cva) *gr->cva = total;
    if (gr->cva_stderr) *gr->cva_stderr = NAN;
    const double inv = 1.0 / (double)P, dt = p->grid.dt;
    double dv[kFactors * kTenors], dm[kTenors];
    for (int k = 0; k < kFactors * kTenors; ++k) dv[k] = sums[k] * inv;
    // drift: dt (A Lambda + B mu_sum e_0) / P with Lambda the sum over dates and paths of
    // the date-end adjoints and mu_sum = dt sum_e e (sum over paths of beta_e): beta_e
    // feeds mu in the e intervals before date e
    {
        std::vector<double> A((size_t)kTenors * kTenors), B((size_t)kTenors * kTenors);
        adjoint_block_sums((double)step_ratio_f32(cc), p->grid.stride, A.data(), B.data());
        double musum = 0.0;
        for (int e = 1; e <= E; ++e) musum += (double)e * sums[kOut + e];
        musum *= dt;
        const double* L = sums.data() + 3 * kTenors;
        for (int j = 0; j < kTenors; ++j) {
            double a = B[(size_t)j * kTenors] * musum;
            for (int m = 0; m < kTenors; ++m) a += A[(size_t)j * kTenors + m] * L[m];
            dm[j] = dt * a * inv;
        }
    }
    if (gr->d_spot)
        for (int j = 0; j < kTenors; ++j) gr->d_spot[j] = sums[4 * kTenors + j] * inv;
    if (gr->d_vol) memcpy(gr->d_vol, dv, sizeof(dv));
    if (gr->d_drift) memcpy(gr->d_drift, dm, sizeof(dm));
    if (gr->d_vol_total) {
        memcpy(gr->d_vol_total, dv, sizeof(dv));
        drift_vjp(market->volatilities, portfolio->dtau, dm, gr->d_vol_total);
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

int hjm_portfolio_autotune(const hjm_market* market, const hjm_portfolio* portfolio,
                           const hjm_config* config, hjm_tuning* best, double* best_ms)
{
    if (!best || !config || !portfolio) return HJM_ERR_NULL;
    hjm_config cfg = *config;
    if (cfg.rng == HJM_RNG_HOST) cfg.rng = HJM_RNG_PHILOX;
    // the VALU path kernel's cost is linear in the horizon: it is tuned on two dates
    hjm_portfolio short_pf = *portfolio;
    std::vector<int32_t> per;
    // the plan's own choice of path kernel (pplan_init)
    const bool mfma = cfg.precision == HJM_FP32_KAHAN && cfg.path_kernel != HJM_KERNEL_VECTOR &&
                      cfg.dt > 0.0 && llround(portfolio->dtau / cfg.dt) == kMfmaStride;
    int nd = 0;
    HJM_TRY(check_portfolio(portfolio, &nd));
    if (!mfma) {
        short_pf.horizon = 2.0 * portfolio->dtau;
        per.assign(portfolio->periods, portfolio->periods + portfolio->nswaps);
        for (auto& x : per) x = std::min(x, 2);
        short_pf.periods = per.data();
        cfg.pfe_quantile = 0.0;
    }
    hjm_portfolio_plan p = nullptr;
    HJM_TRY(hjm_portfolio_plan_create(&p, market, &short_pf, &cfg, nullptr));
    DeviceGuard guard;
    int rc = guard.set(p->device);
    // mean of 3 eager executes after one untimed, per timing slot
    auto measure = [&](const hjm_tuning& t, double (&ms)[5]) {
        pplan_apply_tuning(p, t);
        for (double& m : ms) m = 0.0;
        double tm[5];
        hjm_portfolio_results r{};
        r.timings_ms = tm;
        for (int rep = 0; rep < 4 && rc == HJM_OK; ++rep) {
            rc = pplan_enqueue(p);
            if (rc == HJM_OK) rc = pplan_collect(p, &r);
            if (rep > 0)
                for (int i = 0; i < 5; ++i) ms[i] += tm[i] / 3.0;
        }
    };
    hjm_tuning bt{256, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    double ms[5];
    if (p->use_mfma) {
        // path kernel: shape x {one workgroup per group, persistent (one per CU)}
        const int shapes[4][2] = {{2, 8}, {2, 12}, {1, 12}, {1, 16}};
        double bv = 1e300;
        hjm_tuning t = bt;
        for (const auto& sh : shapes)
            for (int bpc : {0, 1}) {
                t.path_tiles = sh[0];
                t.path_waves = sh[1];
                t.blocks_per_cu = bpc;
                measure(t, ms);
                if (rc == HJM_OK && ms[0] < bv) {
                    bv = ms[0];
                    bt.path_tiles = sh[0];
                    bt.path_waves = sh[1];
                    bt.blocks_per_cu = bpc;
                }
            }
        // valuation: tiles x waves x LDS stage (stages below the M-tile count only)
        const int mtiles = value_mtiles(p->nsets);
        bv = 1e300;
        t = bt;
        for (int tiles : {4, 2})
            for (int waves : {4, 8})
                for (int stage : {8, 4, 2}) {
                    if (stage != 8 && stage >= mtiles) continue;
                    t.value_tiles = tiles;
                    t.value_waves = waves;
                    t.value_stage = stage;
                    measure(t, ms);
                    if (rc == HJM_OK && ms[1] < bv) {
                        bv = ms[1];
                        bt.value_tiles = tiles;
                        bt.value_waves = waves;
                        bt.value_stage = stage;
                    }
                }
        // the exposure pass against the fused valuation-statistics pass (which keeps its
        // own shape): everything after the path kernel
        double passes[2] = {1e300, 1e300};
        for (int fuse : {0, 1}) {
            t = bt;
            t.fuse_statistics = fuse;
            measure(t, ms);
            if (rc == HJM_OK) passes[fuse] = ms[4] - ms[0];
        }
        bt.fuse_statistics = passes[1] < passes[0] ? 1 : 0;
    } else if (!p->use_mfma64) {
        // (the FP64 matrix-core path kernel has one launch shape: nothing to sweep)
        double bv = 1e300;
        for (int bs : {64, 128, 256})
            for (int bpc : {0, 1, 2, 4, 8}) {
                hjm_tuning t = bt;
                t.block_size = bs;
                t.blocks_per_cu = bpc;
                measure(t, ms);
                if (rc == HJM_OK && ms[0] < bv) {
                    bv = ms[0];
                    bt.block_size = bs;
                    bt.blocks_per_cu = bpc;
                }
            }
    }
    double total = 0.0;
    if (rc == HJM_OK) {
        measure(bt, ms);
        total = ms[4];
    }
    hjm_portfolio_plan_destroy(p);
    if (rc != HJM_OK) return rc;
    bt.use_graph = 1;
    *best = bt;
    if (best_ms) *best_ms = total;
    return HJM_OK;
}

}  // extern "C"
