// hjm_adjoint_kernels.hip -- pathwise adjoint of the reference-semantics CVA, gfx950.
//
// Per path J = sum_t w_t max(V_t, 0); its sensitivities to the initial curve (51), the
// loadings (153) and the drift table (51) come from one backward sweep over the Euler
// steps. The step Jacobian T = I + (dt/dtau) Dfwd does not depend on the state, so
// the sweep needs no forward states between exposure dates: the forward kernel's
// numeraire pairs at the dates (one per 50 steps) are the only checkpoint, and the
// factor shocks are read back from the forward pass or regenerated in reverse from the
// counter-based generator.
//
// Backward recursion (n = N..1, injections at the exposure dates n_i = (i-1) stride):
//   mu(n) = mu(n+1) + gA_i e_i,   lambda(n) = T^T lambda(n+1) + mu(n) + gx_i e_i
//   dJ/dsigma_kj += lambda_j(n) sqrt(dt) phi_k(n),  dJ/dm_j += dt lambda_j(n)
// Two lanes share a path: the lower half-wave holds tenors 0..25, the upper half tenors
// 25..50 (its tenor 25 a copy received from the lower lane every step: lambda_26 reads
// lambda_25, and no lower tenor reads an upper one), so lambda, mu and the three loading
// gradients take 130 registers per lane and the recurrence runs once. The drift gradient
// needs no per-step work: within a date block mu is constant, so
//   sum over the block's steps of lambda(n) = A lambda_e + B mu,
//   A = sum_{k<S} (T^T)^k,  B = sum_{r<S-1} (S-1-r) (T^T)^r,
// with lambda_e the adjoint at the block's end; the kernel sums lambda_e over blocks and
// paths and the host applies A to it and B to the mu injections (their sums over paths,
// weighted by the number of blocks each is live in).
#include "hjm_adjoint_lanes.h"
#include "hjm_common.h"
#include "hjm_kernels.h"
#include "hjm_rng.h"
#include "hjm_valu.h"

namespace hjm {

// Seeds of the sweep from the forward pass's numeraire pairs (FP64): CF_i, V_t, the
// survival-weighted indicator sums u_i and gx_i, gA_i, stored [slot][path] in FP32; and
// per workgroup the sums over its paths of gA at every slot, sa_partials[block][51]
// (slot 0 and slots past the last date 0), for the drift gradient's mu terms.
__global__ __launch_bounds__(kBlock) void ref_adjoint_seed_kernel(const ExposureParams e,
                                                                  double dt,
                                                                  const double* __restrict__ nx,
                                                                  const double* __restrict__ ny,
                                                                  float* __restrict__ gx,
                                                                  float* __restrict__ gA,
                                                                  double* __restrict__ sa_partials)
{
    __shared__ double red[kBlock / kWave][kTenors];
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const bool active = i < e.paths;
    const int64_t P = e.paths, ic = active ? i : 0;
    const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
    // V_t = sum_{i > t} CF_i, built backwards; u_i needs V_0..V_{i-1}, so first collect
    // the signs of V_t in a bit mask (D <= 50 dates).
    uint64_t pos = 0;
    double V = 0.0;
    for (int t = e.ndates - 1; t >= 0; --t) {
        const int s = t + 1;
        const double x = nx[(int64_t)s * P + ic], y = ny[(int64_t)s * P + ic];
        V += y * e.notional * e.accrual[s] * (e.m * (exp(x / e.m) - 1.0) - e.K);
        if (V > 0.0) pos |= (1ull << t);
    }
    double u = 0.0;
    for (int s = 1; s <= e.ndates; ++s) {
        if (pos & (1ull << (s - 1))) u += e.w[s - 1];
        const double x = nx[(int64_t)s * P + ic], y = ny[(int64_t)s * P + ic];
        const double base = y * e.notional * e.accrual[s];
        const double cf = base * (e.m * (exp(x / e.m) - 1.0) - e.K);
        const float a = active ? (float)(-u * dt * cf) : 0.f;
        if (active) {
            gx[(int64_t)s * P + i] = (float)(u * base * exp(x / e.m));
            gA[(int64_t)s * P + i] = a;
        }
        const double ws = wave_sum((double)a);
        if (lane == 0) red[wave][s] = ws;
    }
    if (lane == 0) {
        red[wave][0] = 0.0;
        for (int s = e.ndates + 1; s < kTenors; ++s) red[wave][s] = 0.0;
    }
    __syncthreads();
    if (threadIdx.x < kTenors) {
        double t = 0.0;
#pragma unroll
        for (int w = 0; w < kBlock / kWave; ++w) t += red[w][threadIdx.x];
        sa_partials[(int64_t)blockIdx.x * kTenors + threadIdx.x] = t;
    }
}

struct AdjGrid {
    int stride, nsteps, ndates, pad;
    int64_t paths, path0;
    uint32_t seed_lo, seed_hi;
    float c, omc, opc, sqrt_dt;
};

// kSrc: 0 regenerate the in-kernel normals, 1 host-supplied FP64 [path][step][3],
// 2 the forward pass's stored FP32 normals [step][3][path].
// partials[workgroup][5 * 51]: loading gradients g0, g1, g2 (tenor-major within each),
// sum over blocks of lambda_e, and d J / d spot, each summed over the workgroup's paths.
template <int kSrc>
__global__ __launch_bounds__(kBlock) void ref_adjoint_kernel(const AdjGrid g,
                                                             const double* __restrict__ normals,
                                                             const float* __restrict__ zs,
                                                             const float* __restrict__ gx,
                                                             const float* __restrict__ gA,
                                                             double* __restrict__ partials)
{
    __shared__ double acc[kBlock / kWave][2][kAdjKinds][kHalfSlots];
    __shared__ float lacc[kHalfSlots][kBlock];  // per lane: sum over blocks of lambda_e
    __shared__ float zbuf[kSrc == 0 ? kGroupShocks * kAdjPaths : 1];
    const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
    const int half = lane >> 5;                      // 0: tenors 0..25, 1: tenors 25..50
    const int pslot = wave * 32 + (lane & 31);       // the path's index in the workgroup
    for (int k = threadIdx.x; k < (kBlock / kWave) * 2 * kAdjKinds * kHalfSlots; k += kBlock)
        (&acc[0][0][0][0])[k] = 0.0;
#pragma unroll
    for (int i = 0; i < kHalfSlots; ++i) lacc[i][threadIdx.x] = 0.f;
    __syncthreads();

    const int64_t i = (int64_t)blockIdx.x * kAdjPaths + pslot;
    const bool active = i < g.paths;
    const int64_t P = g.paths, ic = active ? i : 0;
    const uint32_t pid = (uint32_t)(g.path0 + ic);
    float* col = zbuf + (kSrc == 0 ? pslot : 0);
    const bool upper = half == 1;
    // the coefficients of slots 25 and 24: tenor 50 (1 + c, and -c on tenor 49's row) in
    // the upper half, an interior tenor (1 - c, 0) in the lower half
    float cv = g.c, omc = g.omc, a25 = upper ? g.opc : g.omc, b24 = upper ? -g.c : 0.f;
    asm volatile("" : "+v"(cv), "+v"(omc), "+v"(a25), "+v"(b24));
    const int tbase = 25 * half;   // tenor of slot 0

    float lam[kHalfSlots], mu[kHalfSlots], g0[kHalfSlots], g1[kHalfSlots], g2[kHalfSlots];
#pragma unroll
    for (int k = 0; k < kHalfSlots; ++k) lam[k] = mu[k] = g0[k] = g1[k] = g2[k] = 0.f;

    // the half-wave sums of a lane family go to this wave's FP64 accumulators (one lane per
    // half adds, always the same lane: a fixed order)
    auto flush = [&](float (&v)[kHalfSlots], int kind) {
#pragma unroll
        for (int k = 0; k < kHalfSlots; ++k) v[k] = half_wave_sum(v[k]);
        if ((lane & 31) == 31) {
#pragma unroll
            for (int k = 0; k < kHalfSlots; ++k) atomicAdd(&acc[wave][half][kind][k], (double)v[k]);
        }
#pragma unroll
        for (int k = 0; k < kHalfSlots; ++k) v[k] = 0.f;
    };

    // one reverse Euler step with the step's normals
    auto step = [&](float z0, float z1, float z2) {
        const float s0 = g.sqrt_dt * z0, s1 = g.sqrt_dt * z1, s2 = g.sqrt_dt * z2;
#pragma unroll
        for (int k = 0; k < kHalfSlots; ++k) {
            op_fmac_vv(g0[k], lam[k], s0);
            op_fmac_vv(g1[k], lam[k], s1);
            op_fmac_vv(g2[k], lam[k], s2);
        }
        // the upper lanes take tenor 25 from the lower lanes before it changes
        const uint32_t l25 = __builtin_bit_cast(uint32_t, lam[kHalfSlots - 1]);
        const float recv = __builtin_bit_cast(float, __builtin_amdgcn_permlane32_swap(l25, l25, false, false)[0]);
        lam[0] = upper ? recv : lam[0];
        // lambda <- T^T lambda + mu, descending so lambda_{k-1} is still old:
        //   slot 25: a25 l25 + c l24 + mu25;  slot 24: (1-c) l24 + c l23 + b24 l25 + mu24
        //   1 <= k <= 23: (1-c) l_k + c l_{k-1} + mu_k;  slot 0: (1-c) l0 + mu0
        float t25, t24;
        op_fma_vvv(t25, cv, lam[24], mu[25]);
        op_fma_vvv(t24, cv, lam[23], mu[24]);
        op_fma_vvv(t24, b24, lam[25], t24);
        op_fma_inplace(lam[25], a25, t25);
        op_fma_inplace(lam[24], omc, t24);
        constexpr int G = 4;
#pragma unroll
        for (int k0 = 23; k0 >= 1; k0 -= G) {
            float t[G];
#pragma unroll
            for (int u = 0; u < G; ++u)
                if (k0 - u >= 1) op_fma_vvv(t[u], cv, lam[k0 - u - 1], mu[k0 - u]);
#pragma unroll
            for (int u = 0; u < G; ++u)
                if (k0 - u >= 1) op_fma_inplace(lam[k0 - u], omc, t[u]);
        }
        op_fma_inplace(lam[0], omc, mu[0]);
    };

    const int S = g.stride;
    constexpr int kPre = 5;   // stored normals: steps read ahead (the stride is a multiple)
    float zq[kPre][3];
    auto loadz = [&](int n, float (&d)[3]) {
        const float* q = zs + (int64_t)n * 3 * P + ic;
        d[0] = q[0];
        d[1] = q[P];
        d[2] = q[2 * P];
    };
    if constexpr (kSrc == 2) {
#pragma unroll
        for (int u = 0; u < kPre; ++u) loadz((g.ndates - 1) * S + S - 1 - u, zq[u]);
    }
    const int whole = S / kGroupSteps * kGroupSteps;
    for (int b = g.ndates - 1; b >= 1; --b) {
        // exposure date n = b * stride records slot b + 1 (tenor b + 1 >= 2)
        const int slot = b + 1 - tbase;
        const float sx = active ? gx[(int64_t)(b + 1) * P + ic] : 0.f;
        const float sa = active ? gA[(int64_t)(b + 1) * P + ic] : 0.f;
#pragma unroll
        for (int k = 0; k < kHalfSlots; ++k) {
            if (k == slot && !(upper && k == 0)) {
                lam[k] += sx + sa;
                mu[k] += sa;
            }
        }
#pragma unroll
        for (int k = 0; k < kHalfSlots; ++k) lacc[k][threadIdx.x] += lam[k];
        if constexpr (kSrc == 2) {
            for (int s0 = S - 1; s0 >= 0; s0 -= kPre) {
#pragma unroll
                for (int u = 0; u < kPre; ++u) {
                    const float a0 = zq[u][0], a1 = zq[u][1], a2 = zq[u][2];
                    const int nn = b * S + s0 - kPre - u;   // same slot of the next chunk
                    if (nn >= S) loadz(nn, zq[u]);
                    step(a0, a1, a2);
                }
            }
        } else {
            for (int s = S - 1; s >= 0; --s) {
                const int n = b * S + s;  // global step index of this Euler step
                if constexpr (kSrc == 1) {
                    const double* q = normals + (ic * g.nsteps + n) * 3;
                    step((float)q[0], (float)q[1], (float)q[2]);
                } else {
                    // groups of 16 steps staged in LDS, entered from their last step
                    Normals3 z;
                    if (s < whole) {
                        if ((s & (kGroupSteps - 1)) == kGroupSteps - 1) {
                            __syncthreads();   // the previous group's reads are done
                            fill_group_half(g.seed_lo, g.seed_hi, pid, (uint32_t)b, s / kGroupSteps,
                                            half, col, kAdjPaths);
                            __syncthreads();   // both lanes' calls are in the column
                        }
                        z = group_step(col, kAdjPaths, s);
                    } else {
                        z = step_normals(g.seed_lo, g.seed_hi, pid, (uint32_t)n, (uint32_t)S);
                    }
                    step(z.z0, z.z1, z.z2);
                }
            }
        }
        // FP32 partials cover at most two blocks, then go to the FP64 accumulators
        if ((b & 1) == 1 || b == 1) {
            flush(g0, 0);
            flush(g1, 1);
            flush(g2, 2);
        }
    }
    // n = 0: lam = T^T lambda(1) + mu(1); dJ/dspot = lam + (stride - 1) mu(1)
    //        + (gx_1 + stride gA_1) e_1   (block 0 holds the curve for stride steps)
    {
        const float sx = active ? gx[P + ic] : 0.f;
        const float sa = active ? gA[P + ic] : 0.f;
        float d[kHalfSlots], la[kHalfSlots];
#pragma unroll
        for (int k = 0; k < kHalfSlots; ++k) {
            d[k] = __builtin_fmaf((float)(S - 1), mu[k], lam[k]);
            if (k == 1 && !upper) d[k] += sx + (float)S * sa;
            la[k] = lacc[k][threadIdx.x];
        }
        flush(la, 3);
        flush(d, 4);
    }
    __syncthreads();
    // output o = kind * 51 + tenor; tenor 25 comes from the lower half, whose slot 25 it is
    for (int o = threadIdx.x; o < kAdjKinds * kTenors; o += kBlock) {
        const int kind = o / kTenors, t = o % kTenors;
        const int h = t > 25 ? 1 : 0, k = t - 25 * h;
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < kBlock / kWave; ++w) s += acc[w][h][kind][k];
        partials[(int64_t)blockIdx.x * (kAdjKinds * kTenors) + o] = s;
    }
}

hipError_t launch_ref_adjoint_seed(const ExposureParams& e, double dt, const double* nx,
                                   const double* ny, float* gx, float* gA, double* sa_partials,
                                   hipStream_t st)
{
    const dim3 grid((unsigned)((e.paths + kBlock - 1) / kBlock));
    ref_adjoint_seed_kernel<<<grid, kBlock, 0, st>>>(e, dt, nx, ny, gx, gA, sa_partials);
    return hipGetLastError();
}

int64_t ref_adjoint_blocks(int64_t paths) { return (paths + kAdjPaths - 1) / kAdjPaths; }

hipError_t launch_ref_adjoint(const RefGrid& rg, double c, double sqrt_dt, const double* normals,
                              const float* zs, const float* gx, const float* gA, double* partials,
                              hipStream_t st)
{
    AdjGrid g{};
    g.stride = rg.stride;
    g.nsteps = rg.nsteps;
    g.ndates = rg.ndates;
    g.paths = rg.paths;
    g.path0 = rg.path0;
    g.seed_lo = rg.seed_lo;
    g.seed_hi = rg.seed_hi;
    g.c = step_ratio_f32(c);  // the FP32 forward step's coefficients (exact pairs)
    g.omc = 1.0f - g.c;
    g.opc = 1.0f + g.c;
    g.sqrt_dt = (float)sqrt_dt;
    const dim3 grid((unsigned)ref_adjoint_blocks(rg.paths));
    if (zs && rg.stride % 5 != 0) return hipErrorInvalidValue;  // stored-normal read-ahead
    if (normals)
        ref_adjoint_kernel<1><<<grid, kBlock, 0, st>>>(g, normals, zs, gx, gA, partials);
    else if (zs)
        ref_adjoint_kernel<2><<<grid, kBlock, 0, st>>>(g, normals, zs, gx, gA, partials);
    else
        ref_adjoint_kernel<0><<<grid, kBlock, 0, st>>>(g, normals, zs, gx, gA, partials);
    return hipGetLastError();
}

}  // namespace hjm
