// hjm_kernels.h -- launchers of the HJM kernels (internal to libhjm_cva.so).
#pragma once
#include "hjm_common.h"

namespace hjm {

struct SelRowState;  // hjm_select.h

struct ExposureParams {
    int64_t paths;
    int ndates;
    int pad;
    double notional, K;
    double m;                // 1 / dtau (compounding frequency of the simple rate)
    double accrual[kTenors];
    double w[kTenors];       // CVA weights (1-R) x default probability shares
};

hipError_t launch_ref_paths_f64(const RefStepConsts<double>& k, const RefGrid& g,
                                const double* normals, double* nx, double* ny, hipStream_t st);
hipError_t launch_ref_paths_f32k(const RefStepConsts<float>& k, const RefGrid& g,
                                 const double* normals, double* nx, double* ny, hipStream_t st);
hipError_t launch_ref_exposure(const ExposureParams& e, const double* nx, const double* ny,
                               double* exposures, double* partials, hipStream_t st);
hipError_t launch_column_sum(const double* partials, int64_t rows, int cols, double* out,
                             hipStream_t st);
// k-th smallest of each row; ws: workspace of bracket_workspace_bytes (hjm_select.h) or null
hipError_t launch_radix_select(const double* x, int rows, int64_t n, int64_t k, void* ws,
                               double* out, hipStream_t st);
hipError_t launch_normals(uint64_t seed, int64_t path0, int64_t paths, int nsteps, int stride,
                          double* out, hipStream_t st);
hipError_t launch_box_muller_parts(const uint32_t* words, int64_t n, float* r, float* c, float* s,
                                   hipStream_t st);

// Matrix-core FP32 path kernels (50 steps per date); prepare once per plan (sets the
// kernels' dynamic LDS limits) before launching or capturing them.
hipError_t mfma_kernels_prepare();
// zs (optional): the forward's FP32 normals [step][factor][path] for a later reverse sweep
hipError_t launch_ref_paths_mfma(const MfmaGrid& g, const uint16_t* wfrag, double* nx, double* ny,
                                 float* zs, hipStream_t st);
// The FP64 matrix-core path kernel (reference semantics); wfrag: build_mfma64_block_map's
// fragments of ref_mfma_tiles(g.ndates) M-tiles; prepare once per plan.
hipError_t mfma64_kernels_prepare();
hipError_t launch_ref_paths_mfma64(const MfmaGrid& g, const double* wfrag, double* nx, double* ny,
                                   float* zs, hipStream_t st);
// The FP64 portfolio path kernel: curves [ndates][51][paths] and logb [ndates][paths];
// wfrag: build_mfma64_block_map's portfolio fragments.
hipError_t launch_pf_paths_mfma64(const MfmaGrid& g, const double* wfrag, double* curves, double* logb,
                                  float* zs, hipStream_t st);
// Launch shape of a matrix-core portfolio kernel (0 in a field: the kernel's default).
//   tiles: N-tiles of 16 paths per wave; waves: per workgroup;
//   stage (valuation): M-tiles of -gamma staged in LDS at a time;
//   max_blocks (paths): a persistent grid of at most this many workgroups.
struct MfmaShape {
    int tiles = 0, waves = 0, stage = 0;
    int64_t max_blocks = 0;
};
// path kernel: {tiles, waves} in {2, 8} (default), {2, 12}, {1, 12}, {1, 16};
// valuation: tiles 2 or 4 (default), waves 4 (default) or 8, stage 1..8 (default 8)
bool pf_path_shape_ok(int tiles, int waves);
bool pf_value_shape_ok(int tiles, int waves);
hipError_t launch_pf_paths_mfma(const MfmaGrid& g, const uint16_t* wfrag, float* curves,
                                double* logb, float* zs, const MfmaShape& shape, hipStream_t st);

// Pathwise adjoint of the reference-semantics CVA.
// sa_partials: [blocks of 256 paths][51] per-slot sums of gA over the block's paths
hipError_t launch_ref_adjoint_seed(const ExposureParams& e, double dt, const double* nx,
                                   const double* ny, float* gx, float* gA, double* sa_partials,
                                   hipStream_t st);
// One-pass sweep, two lanes per path. normals: host FP64 [path][step][3]; else zs: stored
// FP32 [step][3][path]; else regenerate. partials: [ref_adjoint_blocks(paths)][5 * 51]
// (loading gradients of factors 0, 1, 2, sum of the block-end adjoints, d J / d spot).
int64_t ref_adjoint_blocks(int64_t paths);
hipError_t launch_ref_adjoint(const RefGrid& rg, double c, double sqrt_dt, const double* normals,
                              const float* zs, const float* gx, const float* gA, double* partials,
                              hipStream_t st);

// Portfolio semantics.
hipError_t launch_pf_paths_f64(const RefStepConsts<double>& k, const PfGrid& g, const double* normals,
                               double* curves, double* logb, int block, int grid, hipStream_t st);
hipError_t launch_pf_paths_f32(const RefStepConsts<float>& k, const PfGrid& g, const double* normals,
                               float* curves, double* logb, int block, int grid, hipStream_t st);
hipError_t launch_pf_value_f64(const PfValueParams& v, const double* curves, const double* logb,
                               const double* coef, const double* a, const int* kmax, double* ex,
                               hipStream_t st);
// FP32 plans: the matrix-core valuation over build_value_fragments' tables, of paths
// [0, count) (count < 0: all); filter (or null): only the rows (date, set) whose PFE
// bracket selection is unresolved for rank fk with fcap candidate slots
hipError_t launch_pf_value_mfma(const PfValueParams& v, const float* curves, const double* logb,
                                const uint16_t* gfrag, const float* apad, const int* tilek,
                                float* ex, const MfmaShape& shape, hipStream_t st, int64_t count = -1,
                                const SelRowState* filter = nullptr, int64_t fk = 0,
                                int64_t fcap = 0);
// FP32 plans: valuation, EE and per-path CVA moments (partials
// [set][pf_fused_blocks(paths)][ndates + 2], for the grouped column sums) and, with sel,
// the PFE bracket counting and candidates of every row, in one pass without exposures
// (groups of 128 sets)
int64_t pf_fused_blocks(int64_t paths);
hipError_t launch_pf_value_stats_mfma(const PfValueParams& v, const float* curves, const double* logb,
                                      const uint16_t* gfrag, const float* apad, const int* tilek,
                                      const double* w, double* partials, SelRowState* sel,
                                      float* cand, int64_t cap, int stage_slots, hipStream_t st);
// sums[set][c] = sum over blocks of partials[set][block][c] (fixed order)
hipError_t launch_grouped_column_sum(const double* partials, int nsets, int64_t blocks, int cols,
                                     double* sums, hipStream_t st);
// selws (or null): bracket-selection workspace whose counting this pass does, for rank
// pfe_rank of each (date, set) row, after launch_bracket_sample*; needs
// pf_stats_stage(...) > 0 (LDS staging slice per wave, in entries; *flush_dates, if given,
// receives the number of dates between the slice's flushes)
int pf_stats_stage(int64_t paths, int64_t pfe_rank, int ndates, int elem_bytes,
                   int* flush_dates = nullptr);
// lo_parts (or null; FP32 with selws): [ndates * nsets][pf_stats_blocks(paths)] counts
// below the bracket per row and statistics workgroup, stored plainly and added to the rows'
// states by one small launch after the pass (null, and FP64 rows: one atomic per row and
// workgroup instead)
int64_t pf_stats_blocks(int64_t paths);
hipError_t launch_pf_stats_f64(const PfValueParams& v, const double* ex, const double* w,
                               double* partials, double* sums, void* selws, int64_t pfe_rank,
                               uint32_t* lo_parts, hipStream_t st);
hipError_t launch_pf_stats_f32(const PfValueParams& v, const float* ex, const double* w,
                               double* partials, double* sums, void* selws, int64_t pfe_rank,
                               uint32_t* lo_parts, hipStream_t st);
hipError_t launch_radix_select_f32(const float* x, int rows, int64_t n, int64_t k, void* ws,
                                   double* out, hipStream_t st);
// the sampled-bracket selection split around a statistics pass that does its counting
hipError_t launch_bracket_sample(const double* x, int rows, int64_t n, int64_t k, void* ws,
                                 hipStream_t st);
hipError_t launch_bracket_sample_f32(const float* x, int rows, int64_t n, int64_t k, void* ws,
                                     hipStream_t st);
hipError_t launch_bracket_finish(const double* x, int rows, int64_t n, int64_t k, void* ws,
                                 double* out, hipStream_t st);
hipError_t launch_bracket_finish_f32(const float* x, int rows, int64_t n, int64_t k, void* ws,
                                     double* out, hipStream_t st);

// FP64 plans; beta_partials: [blocks of 256 paths][ndates] sums of beta over the block's paths
hipError_t launch_pf_adjoint_seed_f64(const PfValueParams& v, const double* curves,
                                      const double* logb, const double* coef, const double* a,
                                      const int* kmax, const double* w, float* gbar, float* beta,
                                      double* beta_partials, hipStream_t st);
// FP32 plans: both products on the matrix cores over build_value_fragments' and
// build_adjoint_fragments' tables; beta_partials: [pf_adjoint_seed_mfma_blocks(paths)][ndates]
int64_t pf_adjoint_seed_mfma_blocks(int64_t paths);
hipError_t launch_pf_adjoint_seed_mfma(const PfValueParams& v, const float* curves, const double* logb,
                                       const uint16_t* gfrag, const uint16_t* tfrag, const float* apad,
                                       const int* tilek, const double* w, float* gbar, float* beta,
                                       double* beta_partials, hipStream_t st);
// One-pass sweep, two lanes per path; partials [pf_adjoint_blocks(paths)][5 * 51] as the
// reference sweep's.
int64_t pf_adjoint_blocks(int64_t paths);
hipError_t launch_pf_adjoint(const PfGrid& g, double c, const double* normals, const float* zs,
                             const float* gbar, const float* beta, double* partials, hipStream_t st);

// The reference's own GPU design (baseline form).
hipError_t launch_blockwise_paths(const RefStepConsts<double>& k, const RefGrid& g, int block_index,
                                  const double* normals, const double* state_in, double* state_out,
                                  double* integral, double* nx, double* ny, hipStream_t st);
hipError_t launch_blockwise_exposure(const ExposureParams& e, const double* nx, const double* ny,
                                     double* exposures, hipStream_t st);
hipError_t launch_column_mean_gemv(const double* exposures, int64_t paths, double* sums,
                                   hipStream_t st);

}  // namespace hjm
