// hjm_portfolio_kernels.hip -- batched swap portfolios over netting sets, gfx950.
//
// Portfolio semantics (successor of the reference's single-swap diagonal numeraire,
// simulation_cpu.h:248-250,279-295, which cannot express maturities past the 25y grid):
// dates t_e = e dtau from t = 0; bank-account numeraire from the short rate f(t, 0);
// bond prices from the trapezoid integral of the simulated Musiela curve, flat beyond
// the last tenor; each swap is a par floating leg minus its fixed annuity; exposures are
// netted per set. Three passes:
//   1. pf_paths_kernel: persistent, one path per lane, curve in VGPRs; at every date it
//      checkpoints the curve and the log bank account (coalesced [date][tenor][path]).
//   2. pf_value_kernel: per path and date, 60 bond prices and every netting set's value
//      V_g = a_ge - sum_k gamma_gek P_k (host-aggregated coefficients, read through
//      scalar loads), exposure max(V_g, 0) / B stored [set][date][path].
//   3. pf_stats_kernel: per set, EE numerators and per-path CVA moments (FP64).
#include "hjm_common.h"
#include "hjm_kernels.h"
#include "hjm_rng.h"
#include "hjm_valu.h"

namespace hjm {

template <typename T>
using kconst_ptr_pf = const __attribute__((address_space(4))) T*;

template <typename Real>
__device__ __forceinline__ kconst_ptr_pf<RefStepConsts<Real>> pf_consts_ptr()
{
    return (kconst_ptr_pf<RefStepConsts<Real>>)__builtin_amdgcn_kernarg_segment_ptr();
}

// Euler steps of the whole horizon for one path per lane; the update of each tenor is
// the forward path kernel's five in-place FMAs (no discount integrals are needed here).
template <typename Real, bool kHostNormals>
__global__ __launch_bounds__(kBlock) void pf_paths_kernel(const RefStepConsts<Real> k,
                                                          const PfGrid g,
                                                          const double* __restrict__ normals,
                                                          Real* __restrict__ curves,
                                                          double* __restrict__ logb)
{
    const int64_t P = g.paths;
    Real cv = k.c, mcv = -k.c, omc = k.omc, opc = k.opc;
    asm volatile("" : "+v"(cv), "+v"(mcv), "+v"(omc), "+v"(opc));
    const int64_t ntiles = (P + blockDim.x - 1) / blockDim.x;
    for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int64_t i = tile * blockDim.x + threadIdx.x;
        if (i >= P) continue;
        const uint32_t pid = (uint32_t)(g.path0 + i);
        Real f[kTenors];
#pragma unroll
        for (int j = 0; j < kTenors; ++j) f[j] = k.spot[j];
        double lnB = 0.0;
        for (int e = 0; e <= g.ndates_m1; ++e) {
#pragma unroll
            for (int j = 0; j < kTenors; ++j) curves[((int64_t)e * kTenors + j) * P + i] = f[j];
            logb[(int64_t)e * P + i] = lnB;
            if (e == g.ndates_m1) break;
            for (int s = 0; s < g.stride; ++s) {
                const int n = e * g.stride + s;
                lnB += (double)f[0] * g.dt;
                Real z0, z1, z2;
                if constexpr (kHostNormals) {
                    const double* q = normals + ((int64_t)i * g.nsteps + n) * 3;
                    z0 = (Real)q[0];
                    z1 = (Real)q[1];
                    z2 = (Real)q[2];
                } else {
                    const Normals3 z = step_normals(g.seed_lo, g.seed_hi, pid, (uint32_t)n);
                    z0 = (Real)z.z0;
                    z1 = (Real)z.z1;
                    z2 = (Real)z.z2;
                }
                const Real s0 = k.sqrt_dt * z0, s1 = k.sqrt_dt * z1, s2 = k.sqrt_dt * z2;
                kconst_ptr_pf<RefStepConsts<Real>> kp = pf_consts_ptr<Real>();
                asm volatile("" : "+s"(kp));
                constexpr int L = kTenors - 1;
                Real t50;
                op_fma_vvs(t50, mcv, f[L - 1], kp->ddt[L]);
                constexpr int G = 4;
#pragma unroll
                for (int j0 = 0; j0 < L; j0 += G) {
                    Real t[G];
#pragma unroll
                    for (int u = 0; u < G; ++u)
                        if (j0 + u < L) op_fma_vvs(t[u], cv, f[j0 + u + 1], kp->ddt[j0 + u]);
#pragma unroll
                    for (int u = 0; u < G; ++u)
                        if (j0 + u < L) op_fmac_sv(t[u], kp->v0[j0 + u], s0);
#pragma unroll
                    for (int u = 0; u < G; ++u)
                        if (j0 + u < L) op_fmac_sv(t[u], kp->v1[j0 + u], s1);
#pragma unroll
                    for (int u = 0; u < G; ++u)
                        if (j0 + u < L) op_fmac_sv(t[u], kp->v2[j0 + u], s2);
#pragma unroll
                    for (int u = 0; u < G; ++u)
                        if (j0 + u < L) op_fma_inplace(f[j0 + u], omc, t[u]);
                }
                op_fmac_sv(t50, kp->v0[L], s0);
                op_fmac_sv(t50, kp->v1[L], s1);
                op_fmac_sv(t50, kp->v2[L], s2);
                op_fma_inplace(f[L], opc, t50);
            }
        }
    }
}

// Netting-set exposures at every date for one path per lane (dates uniform per wave).
// coef: [date][set][kpad] gamma (k = 1..kpad stored at index k-1), amax: [date][set] a,
// kmax: [date][set] the last non-zero k (0 = no live swap).
template <typename Real>
__global__ __launch_bounds__(kBlock) void pf_value_kernel(const PfValueParams v,
                                                          const Real* __restrict__ curves,
                                                          const double* __restrict__ logb,
                                                          const Real* __restrict__ coef,
                                                          const Real* __restrict__ aconst,
                                                          const int* __restrict__ kmax,
                                                          Real* __restrict__ exposures)
{
    const int64_t P = v.paths;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int e = blockIdx.y;
    if (i >= P) return;
    Real f[kTenors];
#pragma unroll
    for (int j = 0; j < kTenors; ++j) f[j] = curves[((int64_t)e * kTenors + j) * P + i];
    const double disc = exp(-logb[(int64_t)e * P + i]);
    constexpr int KP = kMaxPeriods;
    Real Pk[KP];
    Real I = (Real)0;
    const Real hw = (Real)(0.5 * v.dtau);
#pragma unroll
    for (int kk = 1; kk <= KP; ++kk) {
        if (kk <= kTenors - 1) I += hw * (f[kk - 1] + f[kk]);
        else I += (Real)v.dtau * f[kTenors - 1];
        Pk[kk - 1] = exp(-I);
    }
    kconst_ptr_pf<Real> cp = (kconst_ptr_pf<Real>)coef;
    kconst_ptr_pf<Real> ap = (kconst_ptr_pf<Real>)aconst;
    kconst_ptr_pf<int> kp = (kconst_ptr_pf<int>)kmax;
    for (int gset = 0; gset < v.nsets; ++gset) {
        const int64_t row = (int64_t)e * v.nsets + gset;
        const int K = kp[row];
        kconst_ptr_pf<Real> gam = cp + row * KP;
        Real V = ap[row];
        // chunks of 4 periods up to the set's last live period; the guard is uniform
        // and the indices into Pk are constants, so Pk stays in VGPRs
#pragma unroll
        for (int k0 = 0; k0 < KP; k0 += 4) {
            if (k0 < K) {
#pragma unroll
                for (int u = 0; u < 4; ++u) V -= gam[k0 + u] * Pk[k0 + u];
            }
        }
        const Real ex = V > (Real)0 ? (Real)((double)V * disc) : (Real)0;
        exposures[(row)*P + i] = ex;
    }
}

// Per set g (blockIdx.y) and 256 paths (blockIdx.x): EE numerators at every date and the
// per-path CVA sample's sum and square sum; fixed-order reductions.
template <typename Real>
__global__ __launch_bounds__(kBlock) void pf_stats_kernel(const PfValueParams v,
                                                          const Real* __restrict__ exposures,
                                                          const double* __restrict__ w,
                                                          double* __restrict__ partials)
{
    __shared__ double red[kBlock / kWave][kMaxDates + 2];
    const int gset = blockIdx.y;
    const int64_t P = v.paths;
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const bool active = i < P;
    const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
    const int nd = v.ndates;
    double cva = 0.0;
    for (int e = 0; e < nd; ++e) {
        const double x = active ? (double)exposures[((int64_t)e * v.nsets + gset) * P + i] : 0.0;
        cva += w[(int64_t)gset * nd + e] * x;
        const double s = wave_sum(x);
        if (lane == 0) red[wave][e] = s;
    }
    const double c1 = wave_sum(cva), c2 = wave_sum(cva * cva);
    if (lane == 0) {
        red[wave][nd] = c1;
        red[wave][nd + 1] = c2;
    }
    __syncthreads();
    const int ncol = nd + 2;
    for (int c = threadIdx.x; c < ncol; c += kBlock) {
        double s = 0.0;
#pragma unroll
        for (int ww = 0; ww < kBlock / kWave; ++ww) s += red[ww][c];
        partials[((int64_t)gset * gridDim.x + blockIdx.x) * ncol + c] = s;
    }
}

// out[g][c] = sum_b partials[g][b][c] in block order.
__global__ __launch_bounds__(kBlock) void grouped_column_sum_kernel(const double* __restrict__ partials,
                                                                    int64_t rows, int cols,
                                                                    double* __restrict__ out)
{
    __shared__ double red[kBlock / kWave];
    const int c = blockIdx.x, gset = blockIdx.y;
    const double* base = partials + (int64_t)gset * rows * cols;
    double s = 0.0;
    for (int64_t r = threadIdx.x; r < rows; r += kBlock) s += base[r * cols + c];
    s = wave_sum(s);
    if (threadIdx.x % kWave == 0) red[threadIdx.x / kWave] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
#pragma unroll
        for (int ww = 0; ww < kBlock / kWave; ++ww) t += red[ww];
        out[(int64_t)gset * cols + c] = t;
    }
}

// k-th smallest of each row of [rows][n] non-negative FP32 values: MSD radix selection
// on the bit patterns (order-preserving for x >= 0), passes of 11/11/10 bits.
__global__ __launch_bounds__(1024) void radix_select_f32_kernel(const float* __restrict__ x,
                                                                int64_t n, int64_t k,
                                                                double* __restrict__ out)
{
    constexpr int kBins = 2048;
    __shared__ uint32_t hist[kBins];
    __shared__ uint32_t s_prefix;
    __shared__ int64_t s_k;
    const float* row = x + (int64_t)blockIdx.x * n;
    if (threadIdx.x == 0) {
        s_prefix = 0;
        s_k = k;
    }
    int shift = 32;
    for (int pass = 0; pass < 3; ++pass) {
        const int width = (pass < 2) ? 11 : 10;
        shift -= width;
        for (int b = threadIdx.x; b < kBins; b += blockDim.x) hist[b] = 0;
        __syncthreads();
        const uint32_t prefix = s_prefix;
        const int hi = shift + width;
        for (int64_t t = threadIdx.x; t < n; t += blockDim.x) {
            const uint32_t bits = __builtin_bit_cast(uint32_t, row[t]);
            const bool match = (hi >= 32) ? true : ((bits >> hi) == (prefix >> hi));
            if (match) atomicAdd(&hist[(bits >> shift) & ((1u << width) - 1u)], 1u);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            int64_t kk = s_k;
            int d = 0;
            for (; d < (1 << width); ++d) {
                if ((int64_t)hist[d] >= kk) break;
                kk -= hist[d];
            }
            s_prefix = prefix | ((uint32_t)d << shift);
            s_k = kk;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = (double)__builtin_bit_cast(float, s_prefix);
}

// ---- launchers -------------------------------------------------------------------
template <typename Real>
static hipError_t pf_paths_launch(const RefStepConsts<Real>& k, const PfGrid& g, const double* normals,
                                  Real* curves, double* logb, int block, int grid, hipStream_t st)
{
    if (block != 64 && block != 128 && block != 256) return hipErrorInvalidValue;
    if (grid <= 0) return hipErrorInvalidValue;
    if (normals)
        pf_paths_kernel<Real, true><<<grid, block, 0, st>>>(k, g, normals, curves, logb);
    else
        pf_paths_kernel<Real, false><<<grid, block, 0, st>>>(k, g, normals, curves, logb);
    return hipGetLastError();
}

hipError_t launch_pf_paths_f64(const RefStepConsts<double>& k, const PfGrid& g, const double* normals,
                               double* curves, double* logb, int block, int grid, hipStream_t st)
{
    return pf_paths_launch<double>(k, g, normals, curves, logb, block, grid, st);
}

hipError_t launch_pf_paths_f32(const RefStepConsts<float>& k, const PfGrid& g, const double* normals,
                               float* curves, double* logb, int block, int grid, hipStream_t st)
{
    return pf_paths_launch<float>(k, g, normals, curves, logb, block, grid, st);
}

template <typename Real>
static hipError_t pf_value_launch(const PfValueParams& v, const Real* curves, const double* logb,
                                  const Real* coef, const Real* a, const int* kmax, Real* ex,
                                  hipStream_t st)
{
    const dim3 grid((unsigned)((v.paths + kBlock - 1) / kBlock), (unsigned)v.ndates);
    pf_value_kernel<Real><<<grid, kBlock, 0, st>>>(v, curves, logb, coef, a, kmax, ex);
    return hipGetLastError();
}

hipError_t launch_pf_value_f64(const PfValueParams& v, const double* curves, const double* logb,
                               const double* coef, const double* a, const int* kmax, double* ex,
                               hipStream_t st)
{
    return pf_value_launch<double>(v, curves, logb, coef, a, kmax, ex, st);
}

hipError_t launch_pf_value_f32(const PfValueParams& v, const float* curves, const double* logb,
                               const float* coef, const float* a, const int* kmax, float* ex,
                               hipStream_t st)
{
    return pf_value_launch<float>(v, curves, logb, coef, a, kmax, ex, st);
}

template <typename Real>
static hipError_t pf_stats_launch(const PfValueParams& v, const Real* ex, const double* w,
                                  double* partials, double* sums, hipStream_t st)
{
    const unsigned nb = (unsigned)((v.paths + kBlock - 1) / kBlock);
    pf_stats_kernel<Real><<<dim3(nb, (unsigned)v.nsets), kBlock, 0, st>>>(v, ex, w, partials);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return err;
    grouped_column_sum_kernel<<<dim3((unsigned)(v.ndates + 2), (unsigned)v.nsets), kBlock, 0, st>>>(
        partials, nb, v.ndates + 2, sums);
    return hipGetLastError();
}

hipError_t launch_pf_stats_f64(const PfValueParams& v, const double* ex, const double* w,
                               double* partials, double* sums, hipStream_t st)
{
    return pf_stats_launch<double>(v, ex, w, partials, sums, st);
}

hipError_t launch_pf_stats_f32(const PfValueParams& v, const float* ex, const double* w,
                               double* partials, double* sums, hipStream_t st)
{
    return pf_stats_launch<float>(v, ex, w, partials, sums, st);
}

hipError_t launch_radix_select_f32(const float* x, int rows, int64_t n, int64_t k, double* out,
                                   hipStream_t st)
{
    if (rows <= 0) return hipSuccess;
    radix_select_f32_kernel<<<rows, 1024, 0, st>>>(x, n, k, out);
    return hipGetLastError();
}

}  // namespace hjm
