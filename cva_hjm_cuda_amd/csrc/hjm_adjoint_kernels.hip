// hjm_adjoint_kernels.hip -- pathwise adjoint of the reference-semantics CVA, gfx950.
//
// Per path J = sum_t w_t max(V_t, 0); its sensitivities to the initial curve (51), the
// loadings (153) and the drift table (51) come from one backward sweep over the Euler
// steps. The step Jacobian T = I + (dt/dtau) Dfwd does not depend on the state, so
// the sweep needs no forward states between exposure dates: the forward kernel's
// numeraire pairs at the dates (one per 50 steps) are the only checkpoint, and the
// factor shocks are regenerated in reverse from the counter-based generator.
//
// Backward recursion (n = N..1, injections at the exposure dates n_i = (i-1) stride):
//   mu(n) = mu(n+1) + gA_i e_i,   lambda(n) = T^T lambda(n+1) + mu(n) + gx_i e_i
//   dJ/dsigma_kj += lambda_j(n) sqrt(dt) phi_k(n),  dJ/dm_j += dt lambda_j(n)
// The sweep runs twice per path (pass 0: factors 0, 1; pass 1: factor 2, drift and the
// curve), which keeps the per-lane state at 204 FP32 registers. Per-lane FP32 partials
// cover one 50-step block, then are reduced across the wave and accumulated in FP64.
#include "hjm_common.h"
#include "hjm_kernels.h"
#include "hjm_rng.h"
#include "hjm_valu.h"

namespace hjm {

// Seeds of the sweep from the forward pass's numeraire pairs (FP64): CF_i, V_t, the
// survival-weighted indicator sums u_i and gx_i, gA_i, stored [slot][path] in FP32.
__global__ __launch_bounds__(kBlock) void ref_adjoint_seed_kernel(const ExposureParams e,
                                                                  double dt,
                                                                  const double* __restrict__ nx,
                                                                  const double* __restrict__ ny,
                                                                  float* __restrict__ gx,
                                                                  float* __restrict__ gA)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= e.paths) return;
    const int64_t P = e.paths;
    // V_t = sum_{i > t} CF_i, built backwards; u_i needs V_0..V_{i-1}, so first collect
    // the signs of V_t in a bit mask (D <= 50 dates).
    uint64_t pos = 0;
    double V = 0.0;
    for (int t = e.ndates - 1; t >= 0; --t) {
        const int s = t + 1;
        const double x = nx[(int64_t)s * P + i], y = ny[(int64_t)s * P + i];
        V += y * e.notional * e.accrual[s] * (e.m * (exp(x / e.m) - 1.0) - e.K);
        if (V > 0.0) pos |= (1ull << t);
    }
    double u = 0.0;
    for (int s = 1; s <= e.ndates; ++s) {
        if (pos & (1ull << (s - 1))) u += e.w[s - 1];
        const double x = nx[(int64_t)s * P + i], y = ny[(int64_t)s * P + i];
        const double base = y * e.notional * e.accrual[s];
        const double cf = base * (e.m * (exp(x / e.m) - 1.0) - e.K);
        gx[(int64_t)s * P + i] = (float)(u * base * exp(x / e.m));
        gA[(int64_t)s * P + i] = (float)(-u * dt * cf);
    }
}

struct AdjGrid {
    int stride, nsteps, ndates, pass;
    int64_t paths, path0;
    uint32_t seed_lo, seed_hi;
    float c, omc, opc, sqrt_dt;
};

// Sum of a lane value over the wave in FP64, fixed butterfly order.
__device__ __forceinline__ double wave_sum_f(float v) { return wave_sum((double)v); }

// kSrc: 0 regenerate the in-kernel normals, 1 host-supplied FP64 [path][step][3],
// 2 the forward pass's stored FP32 normals [step][3][path].
template <int kPass, int kSrc>
__global__ __launch_bounds__(kBlock) void ref_adjoint_sweep_kernel(const AdjGrid g,
                                                                   const double* __restrict__ normals,
                                                                   const float* __restrict__ zs,
                                                                   const float* __restrict__ gx,
                                                                   const float* __restrict__ gA,
                                                                   double* __restrict__ partials)
{
    // per wave FP64 accumulators: pass 0 -> g0, g1 (102); pass 1 -> g2, gm, dspot (153)
    constexpr int kAcc = 3 * kTenors;
    __shared__ double acc[kBlock / kWave][kAcc];
    // each lane's column of regenerated shocks (16 steps at a time)
    __shared__ float zbuf[kSrc == 0 ? kGroupShocks * kBlock : 1];
    float* col = zbuf + (kSrc == 0 ? threadIdx.x : 0);
    const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
    for (int k = lane; k < kAcc; k += kWave) acc[wave][k] = 0.0;
    __syncthreads();

    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const bool active = i < g.paths;
    const int64_t P = g.paths;
    const uint32_t pid = (uint32_t)(g.path0 + (active ? i : 0));
    float cv = g.c, mcv = -g.c, omc = g.omc, opc = g.opc;
    asm volatile("" : "+v"(cv), "+v"(mcv), "+v"(omc), "+v"(opc));

    float lam[kTenors], mu[kTenors], ga[kTenors], gb[kTenors];
#pragma unroll
    for (int j = 0; j < kTenors; ++j) lam[j] = mu[j] = ga[j] = gb[j] = 0.f;

    for (int b = g.ndates - 1; b >= 1; --b) {
        // exposure date n = b * stride records slot b + 1
        const int slot = b + 1;
        const float sx = active ? gx[(int64_t)slot * P + i] : 0.f;
        const float sa = active ? gA[(int64_t)slot * P + i] : 0.f;
#pragma unroll
        for (int j = 1; j < kTenors; ++j) {
            if (j == slot) {
                lam[j] += sx + sa;
                mu[j] += sa;
            }
        }
        // one reverse Euler step with the step's normals
        auto step = [&](float z0, float z1, float z2) {
            (void)z0;
            (void)z1;
            (void)z2;
            if constexpr (kPass == 0) {
                const float s0 = g.sqrt_dt * z0, s1 = g.sqrt_dt * z1;
#pragma unroll
                for (int j = 0; j < kTenors; ++j) {
                    op_fmac_vv(ga[j], lam[j], s0);
                    op_fmac_vv(gb[j], lam[j], s1);
                }
            } else {
                const float s2 = g.sqrt_dt * z2;
#pragma unroll
                for (int j = 0; j < kTenors; ++j) {
                    op_fmac_vv(ga[j], lam[j], s2);
                    op_add_inplace(gb[j], lam[j]);
                }
            }
            // lambda <- T^T lambda + mu, descending so lambda_{j-1} is still old:
            //   j=50: (1+c) l50 + c l49 + mu50;  j=49: (1-c) l49 + c l48 - c l50 + mu49
            //   1<=j<=48: (1-c) l_j + c l_{j-1} + mu_j;  j=0: (1-c) l0 + mu0
            float t50, t49;
            op_fma_vvv(t50, cv, lam[49], mu[50]);
            op_fma_vvv(t49, cv, lam[48], mu[49]);
            op_fma_vvv(t49, mcv, lam[50], t49);
            op_fma_inplace(lam[50], opc, t50);
            op_fma_inplace(lam[49], omc, t49);
            constexpr int G = 4;
#pragma unroll
            for (int j0 = 48; j0 >= 1; j0 -= G) {
                float t[G];
#pragma unroll
                for (int u = 0; u < G; ++u)
                    if (j0 - u >= 1) op_fma_vvv(t[u], cv, lam[j0 - u - 1], mu[j0 - u]);
#pragma unroll
                for (int u = 0; u < G; ++u)
                    if (j0 - u >= 1) op_fma_inplace(lam[j0 - u], omc, t[u]);
            }
            op_fma_inplace(lam[0], omc, mu[0]);
        };
        if constexpr (kSrc == 2) {
            // stored normals, read kPre steps ahead so the HBM latency hides behind the
            // previous steps' arithmetic (the matrix-core forward runs 50 steps per date)
            constexpr int kPre = 10;
            float cur[kPre][2], nxt[kPre][2];
            auto load = [&](int n_hi, float (&d)[kPre][2]) {
#pragma unroll
                for (int u = 0; u < kPre; ++u) {
                    const float* q = zs + (int64_t)(n_hi - u) * 3 * P + (active ? i : 0);
                    if constexpr (kPass == 0) {
                        d[u][0] = q[0];
                        d[u][1] = q[P];
                    } else {
                        d[u][0] = q[2 * P];
                        d[u][1] = 0.f;
                    }
                }
            };
            load(b * g.stride + g.stride - 1, cur);
            for (int s0 = g.stride - 1; s0 >= 0; s0 -= kPre) {
                if (s0 - kPre >= 0) load(b * g.stride + s0 - kPre, nxt);
#pragma unroll
                for (int u = 0; u < kPre; ++u) {
                    if constexpr (kPass == 0) step(cur[u][0], cur[u][1], 0.f);
                    else step(0.f, 0.f, cur[u][0]);
                }
#pragma unroll
                for (int u = 0; u < kPre; ++u) {
                    cur[u][0] = nxt[u][0];
                    cur[u][1] = nxt[u][1];
                }
            }
        } else {
            const int whole = g.stride / kGroupSteps * kGroupSteps;
            for (int s = g.stride - 1; s >= 0; --s) {
                const int n = b * g.stride + s;  // global step index of this Euler step
                if constexpr (kSrc == 1) {
                    const double* q = normals + ((int64_t)(active ? i : 0) * g.nsteps + n) * 3;
                    step((float)q[0], (float)q[1], (float)q[2]);
                } else {
                    // groups of 16 steps staged in LDS, entered from their last step
                    Normals3 z;
                    if (s < whole) {
                        if ((s & (kGroupSteps - 1)) == kGroupSteps - 1)
                            fill_group(g.seed_lo, g.seed_hi, pid, (uint32_t)b, s / kGroupSteps, col, kBlock);
                        z = group_step(col, kBlock, s);
                    } else {
                        z = step_normals(g.seed_lo, g.seed_hi, pid, (uint32_t)n, (uint32_t)g.stride);
                    }
                    step(z.z0, z.z1, z.z2);
                }
            }
        }
        // flush the block's FP32 partials into the wave's FP64 accumulators
#pragma unroll
        for (int j = 0; j < kTenors; ++j) {
            const double a = wave_sum_f(ga[j]), bb = wave_sum_f(gb[j]);
            if (lane == 0) {
                acc[wave][j] += a;
                acc[wave][kTenors + j] += bb;
            }
            ga[j] = gb[j] = 0.f;
        }
    }
    if constexpr (kPass == 1) {
        // n = 0: lam = T^T lambda(1) + mu(1); dJ/dspot = lam + (stride - 1) mu(1)
        //        + (gx_1 + stride gA_1) e_1   (block 0 holds the curve for stride steps)
        const float sx = active ? gx[P + i] : 0.f;
        const float sa = active ? gA[P + i] : 0.f;
#pragma unroll
        for (int j = 0; j < kTenors; ++j) {
            float d = __builtin_fmaf((float)(g.stride - 1), mu[j], lam[j]);
            if (j == 1) d += sx + (float)g.stride * sa;
            const double r = wave_sum_f(d);
            if (lane == 0) acc[wave][2 * kTenors + j] += r;
        }
    }
    __syncthreads();
    constexpr int nacc = (kPass == 0) ? 2 * kTenors : 3 * kTenors;
    for (int k = threadIdx.x; k < nacc; k += kBlock) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < kBlock / kWave; ++w) s += acc[w][k];
        partials[(int64_t)blockIdx.x * nacc + k] = s;
    }
}

hipError_t launch_ref_adjoint_seed(const ExposureParams& e, double dt, const double* nx,
                                   const double* ny, float* gx, float* gA, hipStream_t st)
{
    const dim3 grid((unsigned)((e.paths + kBlock - 1) / kBlock));
    ref_adjoint_seed_kernel<<<grid, kBlock, 0, st>>>(e, dt, nx, ny, gx, gA);
    return hipGetLastError();
}

hipError_t launch_ref_adjoint_sweep(const RefGrid& rg, double c, double sqrt_dt, int pass,
                                    const double* normals, const float* zs, const float* gx,
                                    const float* gA, double* partials, hipStream_t st)
{
    AdjGrid g{};
    g.stride = rg.stride;
    g.nsteps = rg.nsteps;
    g.ndates = rg.ndates;
    g.pass = pass;
    g.paths = rg.paths;
    g.path0 = rg.path0;
    g.seed_lo = rg.seed_lo;
    g.seed_hi = rg.seed_hi;
    g.c = step_ratio_f32(c);  // the FP32 forward step's coefficients (exact pairs)
    g.omc = 1.0f - g.c;
    g.opc = 1.0f + g.c;
    g.sqrt_dt = (float)sqrt_dt;
    const dim3 grid((unsigned)((rg.paths + kBlock - 1) / kBlock));
    if (zs && rg.stride % 10 != 0) return hipErrorInvalidValue;  // stored-normal prefetch chunks
#define HJM_SWEEP(P_, S_) \
    ref_adjoint_sweep_kernel<P_, S_><<<grid, kBlock, 0, st>>>(g, normals, zs, gx, gA, partials)
    const int src = normals ? 1 : (zs ? 2 : 0);
    if (pass == 0) {
        if (src == 1) HJM_SWEEP(0, 1);
        else if (src == 2) HJM_SWEEP(0, 2);
        else HJM_SWEEP(0, 0);
    } else {
        if (src == 1) HJM_SWEEP(1, 1);
        else if (src == 2) HJM_SWEEP(1, 2);
        else HJM_SWEEP(1, 0);
    }
#undef HJM_SWEEP
    return hipGetLastError();
}

}  // namespace hjm
