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
//   2. pf_value_kernel (FP64; FP32 plans use pf_value_mfma_kernel, the same values with
//      the set sums on the matrix cores, in hjm_mfma_kernels.hip): per path and date, 60
//      bond prices and every netting set's value V_g = a_ge - sum_k gamma_gek P_k
//      (host-aggregated coefficients staged in LDS), exposure max(V_g, 0) / B stored
//      [date][set][path].
//   3. pf_stats_kernel: per set, EE numerators and per-path CVA moments (FP64).
#include "hjm_adjoint_lanes.h"
#include "hjm_common.h"
#include "hjm_kernels.h"
#include "hjm_rng.h"
#include "hjm_select.h"
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
    // each lane's column of staged shocks (16 steps at a time; dynamic LDS of
    // kGroupShocks x blockDim floats when the normals are generated here)
    extern __shared__ float pf_zbuf[];
    float* col = pf_zbuf + threadIdx.x;
    const int ld = (int)blockDim.x;
    const int whole = g.stride / kGroupSteps * kGroupSteps;
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
                    Normals3 z;
                    if (s < whole) {
                        if ((s & (kGroupSteps - 1)) == 0)
                            fill_group(g.seed_lo, g.seed_hi, pid, (uint32_t)e, s / kGroupSteps, col, ld);
                        z = group_step(col, ld, s);
                    } else {
                        z = step_normals(g.seed_lo, g.seed_hi, pid, (uint32_t)n, (uint32_t)g.stride);
                    }
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

// Netting-set exposures at every date for one path per lane (one date per workgroup).
// coef: [date][set][kpad] gamma (k = 1..kpad stored at index k-1), amax: [date][set] a,
// kmax: [date][set] the last non-zero k (0 = no live swap). The date's coefficient rows are
// staged in LDS, kValueSets sets at a time, and read there with workgroup-uniform
// addresses (broadcasts).
constexpr int kValueSets = 64;

template <typename Real>
__global__ __launch_bounds__(kBlock) void pf_value_kernel(const PfValueParams v,
                                                          const Real* __restrict__ curves,
                                                          const double* __restrict__ logb,
                                                          const Real* __restrict__ coef,
                                                          const Real* __restrict__ aconst,
                                                          const int* __restrict__ kmax,
                                                          Real* __restrict__ exposures)
{
    constexpr int KP = kMaxPeriods;
    __shared__ Real sg[kValueSets * KP];
    __shared__ Real sa[kValueSets];
    __shared__ int sk[kValueSets];
    const int e = blockIdx.y;
    const int64_t P = v.paths;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = i < P;
    const int64_t ic = live ? i : 0;
    Real f[kTenors];
#pragma unroll
    for (int j = 0; j < kTenors; ++j) f[j] = curves[((int64_t)e * kTenors + j) * P + ic];
    const double disc = exp(-logb[(int64_t)e * P + ic]);
    const Real discr = (Real)disc;  // the discount factor in the exposure's precision
    Real Pk[KP];
    Real I = (Real)0;
    const Real hw = (Real)(0.5 * v.dtau);
#pragma unroll
    for (int kk = 1; kk <= KP; ++kk) {
        if (kk <= kTenors - 1) I += hw * (f[kk - 1] + f[kk]);
        else I += (Real)v.dtau * f[kTenors - 1];
        Pk[kk - 1] = exp(-I);
    }
    for (int g0 = 0; g0 < v.nsets; g0 += kValueSets) {
        const int ng = v.nsets - g0 < kValueSets ? v.nsets - g0 : kValueSets;
        __syncthreads();
        const int64_t base = (int64_t)e * v.nsets + g0;
        for (int t = threadIdx.x; t < ng * KP; t += blockDim.x) sg[t] = coef[base * KP + t];
        for (int t = threadIdx.x; t < ng; t += blockDim.x) {
            sa[t] = aconst[base + t];
            sk[t] = kmax[base + t];
        }
        __syncthreads();
        for (int gg = 0; gg < ng; ++gg) {
            const int K = sk[gg];
            const Real* gam = sg + gg * KP;
            Real V = sa[gg];
            // chunks of 16 periods up to the set's last live period: the guard is uniform,
            // the indices into Pk are constants (Pk stays in VGPRs), and each chunk's LDS
            // reads are issued together ahead of its FMAs (one wait per 16 FMAs)
#pragma unroll
            for (int k0 = 0; k0 < KP; k0 += 16) {
                if (k0 < K) {
                    Real gk[16];
#pragma unroll
                    for (int u = 0; u < 16; ++u)
                        if (k0 + u < KP) gk[u] = gam[k0 + u];
#pragma unroll
                    for (int u = 0; u < 16; ++u)
                        if (k0 + u < KP) V -= gk[u] * Pk[k0 + u];
                }
            }
            const Real ex = V > (Real)0 ? V * discr : (Real)0;
            if (live) exposures[(base + gg) * P + i] = ex;
        }
    }
}

// Per set g (blockIdx.y) and kStatPaths paths (blockIdx.x; each thread takes kStatPerThread
// of them, 256 apart, so one wave reduction serves 8 exposures per lane): EE numerators at
// every date and the per-path CVA sample's sum and square sum; fixed-order reductions.
// kCount: the same pass does the counting of the PFE's sampled-bracket selection for the
// rows (date, set) it reads (hjm_select.h): counts below the bracket / zeros, and the
// entries inside it. The counts are wave ballots. Each wave stages its inside entries in
// its own slice of dynamic LDS (stage entries, in date order, at a wave-uniform offset);
// every `flush` dates (and after the last) one atomic per row claims the workgroup's
// candidate slots of those dates, the waves copy their slices out and start them again.
// A slice that fills up flags the rows it could not hold (kSelOverflow in the row's
// claim), whose selection then runs over the whole row.
constexpr int kStatPerThread = 8;
constexpr int kStatPaths = kBlock * kStatPerThread;

// lane mask of an unsigned compare (LLVM predicates 36: <, 37: <=) made by the compare
// itself: a ballot of a condition the compiler has folded into a subtraction's borrow goes
// through a select and a second compare
template <int kPred>
__device__ __forceinline__ uint64_t lanes_where(uint32_t a, uint32_t b)
{
    return __builtin_amdgcn_uicmp(a, b, kPred);
}
template <int kPred>
__device__ __forceinline__ uint64_t lanes_where(uint64_t a, uint64_t b)
{
    return __builtin_amdgcn_uicmpl(a, b, kPred);
}

template <typename Real, bool kCount>
__device__ __forceinline__ void pf_stats_rows(const PfValueParams v,
                                              const Real* __restrict__ exposures,
                                              const double* __restrict__ w,
                                              double* __restrict__ partials,
                                              SelRowState* __restrict__ sel,
                                              Real* __restrict__ cand, int64_t cap, int stage,
                                              int flush, uint32_t* __restrict__ lo_parts)
{
    typedef typename SelectBits<Real>::U U;
    constexpr int kWaves = kBlock / kWave;
    constexpr int kD = kCount ? kMaxDates : 1;
    __shared__ double red[kWaves][kMaxDates + 2];
    __shared__ uint32_t cnt[2][kWaves][kD];  // below-or-zero and inside counts per wave, date
    __shared__ uint32_t beg[kWaves][kD];     // first staged entry of (wave, date)
    __shared__ uint32_t base[kD];            // claimed candidate slot of the workgroup's rows
    __shared__ U s_lo[kD], s_hi[kD];         // the rows' brackets
    __shared__ double s_w[kMaxDates];        // the set's CVA weights
    __shared__ uint8_t s_live[kMaxDates];    // the set has a live swap at the date
    extern __shared__ __align__(16) unsigned char stage_raw[];
    const int gset = blockIdx.y;
    const int64_t P = v.paths;
    const int64_t i0 = (int64_t)blockIdx.x * kStatPaths + threadIdx.x;
    const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
    uint32_t npad = 0;   // wave-uniform: the wave's (k, lane) slots past the last path
#pragma unroll
    for (int k = 0; k < kStatPerThread; ++k)
        npad += (uint32_t)__popcll(__ballot(i0 + (int64_t)k * kBlock >= P));
    U* slice = reinterpret_cast<U*>(stage_raw) + (size_t)wave * stage;
    uint32_t used = 0;  // wave-uniform
    const int nd = v.ndates;
    double cva[kStatPerThread];
#pragma unroll
    for (int k = 0; k < kStatPerThread; ++k) cva[k] = 0.0;
    // weights and brackets staged in LDS: a per-date global load in the loop would make the
    // wait for it also wait for the next date's exposures already in flight
    for (int e = threadIdx.x; e < nd; e += kBlock) {
        s_w[e] = w[(int64_t)gset * nd + e];
        s_live[e] = v.kmax ? v.kmax[(int64_t)e * v.nsets + gset] > 0 : 1;
        if constexpr (kCount) {
            s_lo[e] = (U)sel[(int64_t)e * v.nsets + gset].lo;
            s_hi[e] = (U)sel[(int64_t)e * v.nsets + gset].hi;
        }
    }
    __syncthreads();
    // the next date's exposures are loaded while this date's are reduced and counted
    const bool full = (int64_t)(blockIdx.x + 1) * kStatPaths <= P;  // no path out of range
    Real xn[kStatPerThread];
    // a matured set's row holds zeros: not read
    auto load = [&](int e, Real* dst) {
        const Real* row = exposures + ((int64_t)e * v.nsets + gset) * P;
        if (!s_live[e]) {
#pragma unroll
            for (int k = 0; k < kStatPerThread; ++k) dst[k] = (Real)0;
        } else if (full) {
#pragma unroll
            for (int k = 0; k < kStatPerThread; ++k)
                dst[k] = __builtin_nontemporal_load(row + i0 + (int64_t)k * kBlock);
        } else {
#pragma unroll
            for (int k = 0; k < kStatPerThread; ++k) {
                const int64_t i = i0 + (int64_t)k * kBlock;
                dst[k] = i < P ? __builtin_nontemporal_load(row + i) : (Real)0;
            }
        }
    };
    // adds the workgroup's counts of dates [e0, e1) to their rows, claims the rows'
    // candidate slots and writes the waves' staged entries there (every wave calls it at
    // the same dates)
    auto flush_slices = [&](int e0, int e1) {
        __syncthreads();
        for (int e = e0 + threadIdx.x; e < e1; e += kBlock) {
            const int64_t rid = (int64_t)e * v.nsets + gset;
            uint32_t nlo = 0, nin = 0, nst = 0;
#pragma unroll
            for (int ww = 0; ww < kWaves; ++ww) {
                nlo += cnt[0][ww][e];
                nin += cnt[1][ww][e];
                nst += beg[ww][e] != ~0u ? cnt[1][ww][e] : 0u;
            }
            // FP32 rows: summed per row by sel_add_row_counts_kernel (FP64 rows keep the
            // atomic: the branch costs them a wave per SIMD)
            if (sizeof(Real) == sizeof(float) && lo_parts)
                lo_parts[rid * gridDim.x + blockIdx.x] = nlo;
            else if (nlo)
                atomicAdd(s_lo[e] ? &sel[rid].below : &sel[rid].zin, nlo);
            const uint32_t b = nin ? sel_count_and_claim(&sel[rid], nin, nst) : 0u;  // nst <= nin
            base[e] = nst ? b : 0u;
        }
        __syncthreads();
        // wave w writes the dates e0 + w, e0 + w + kWaves, ...: all waves' staged entries of
        // a date in wave order, one contiguous run of the row's candidates per store
        // instruction (lane i takes entry i - pre[k] of the slice of wave k)
        const U* slices = reinterpret_cast<const U*>(stage_raw);
        for (int e = e0 + wave; e < e1; e += kWaves) {
            uint32_t pre[kWaves + 1], b0[kWaves];
            pre[0] = 0;
#pragma unroll
            for (int ww = 0; ww < kWaves; ++ww) {
                b0[ww] = beg[ww][e];
                pre[ww + 1] = pre[ww] + (b0[ww] != ~0u ? cnt[1][ww][e] : 0u);
            }
            const uint32_t off = base[e];
            Real* out = cand + ((int64_t)e * v.nsets + gset) * cap;
            for (uint32_t i = lane; i < pre[kWaves]; i += kWave) {
                int k = 0;
#pragma unroll
                for (int ww = 1; ww < kWaves; ++ww) k += i >= pre[ww] ? 1 : 0;
                if ((int64_t)off + i < cap)
                    out[off + i] = __builtin_bit_cast(Real, slices[(size_t)k * stage + b0[k] + (i - pre[k])]);
            }
        }
        __syncthreads();   // the slices are read across waves: restarted only after this
    };
    int e_flushed = 0;
    load(0, xn);
    for (int e = 0; e < nd; ++e) {
        const int64_t rid = (int64_t)e * v.nsets + gset;
        const double we = s_w[e];
        double s = 0.0;
        Real xs[kStatPerThread];
#pragma unroll
        for (int k = 0; k < kStatPerThread; ++k) xs[k] = xn[k];
        if (e + 1 < nd) load(e + 1, xn);
#pragma unroll
        for (int k = 0; k < kStatPerThread; ++k) {
            const double x = (double)xs[k];
            s += x;
            cva[k] += we * x;
        }
        s = wave_sum(s);
        if (lane == 0) red[wave][e] = s;
        if constexpr (kCount) {
            const U lo = s_lo[e], hi = s_hi[e];
            const U lo1 = lo ? lo : (U)1;
            const bool any_in = hi >= lo1;
            const U span = any_in ? (U)(hi - lo1) : (U)0;
            const bool collect = any_in && lo != hi;
            // keys past the last path are 0: below every bracket (lo1 >= 1; npad is taken
            // off again) and inside none (0 - lo1 wraps above every span of finite keys)
            uint32_t nlo = 0, nin = 0;
            uint64_t masks[kStatPerThread];
#pragma unroll
            for (int k = 0; k < kStatPerThread; ++k) {
                const U b = __builtin_bit_cast(U, xs[k]);
                nlo += (uint32_t)__popcll(lanes_where<36>(b, lo1));
                const uint64_t mk = lanes_where<37>((U)(b - lo1), span);
                masks[k] = any_in ? mk : 0ull;
                nin += (uint32_t)__popcll(masks[k]);
            }
            nlo -= npad;
            uint32_t staged = 0;
            if (collect && nin) {
                if (used + nin <= (uint32_t)stage) {
                    uint32_t off = used;
#pragma unroll
                    for (int k = 0; k < kStatPerThread; ++k) {
                        const uint64_t mk = masks[k];
                        if (mk) {   // wave-uniform: skips the 64-path groups holding none
                            const U b = __builtin_bit_cast(U, xs[k]);
                            const uint32_t below = __builtin_amdgcn_mbcnt_hi(
                                (uint32_t)(mk >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mk, 0u));
                            if ((U)(b - lo1) <= span) slice[off + below] = b;
                            off += (uint32_t)__popcll(mk);
                        }
                    }
                    staged = nin;
                } else if (lane == 0) {
                    atomicOr(&sel[rid].claim, kSelOverflow);
                }
            }
            if (lane == 0) {
                cnt[0][wave][e] = nlo;
                cnt[1][wave][e] = nin;
                beg[wave][e] = staged ? used : ~0u;
            }
            used += staged;
            if (e + 1 == nd || e + 1 - e_flushed == flush) {
                flush_slices(e_flushed, e + 1);
                e_flushed = e + 1;
                used = 0;
            }
        }
    }
    double c1 = 0.0, c2 = 0.0;
#pragma unroll
    for (int k = 0; k < kStatPerThread; ++k) {
        c1 += cva[k];
        c2 += cva[k] * cva[k];
    }
    c1 = wave_sum(c1);
    c2 = wave_sum(c2);
    if (lane == 0) {
        red[wave][nd] = c1;
        red[wave][nd + 1] = c2;
    }
    __syncthreads();
    const int ncol = nd + 2;
    for (int c = threadIdx.x; c < ncol; c += kBlock) {
        double t = 0.0;
#pragma unroll
        for (int ww = 0; ww < kWaves; ++ww) t += red[ww][c];
        partials[((int64_t)gset * gridDim.x + blockIdx.x) * ncol + c] = t;
    }
}

template <typename Real, bool kCount>
__global__ __launch_bounds__(kBlock) void pf_stats_kernel(const PfValueParams v,
                                                          const Real* __restrict__ exposures,
                                                          const double* __restrict__ w,
                                                          double* __restrict__ partials,
                                                          SelRowState* __restrict__ sel,
                                                          Real* __restrict__ cand, int64_t cap,
                                                          int stage, int flush,
                                                          uint32_t* __restrict__ lo_parts)
{
    pf_stats_rows<Real, kCount>(v, exposures, w, partials, sel, cand, cap, stage, flush, lo_parts);
}

// FP32 rows with the PFE counting: held to 80 VGPRs for 6 waves per SIMD, which it fits
// without spills (it first spilled 20 bytes per lane there and was still faster: 7.87 ->
// 7.63 ms at 1048576 paths x 61 dates x 128 sets). More waves have not paid: with FP32
// per-path CVA sums and a staging slice of 768 entries (two flushes instead of one, LDS
// for 8 workgroups) it took 6.9-7.0 ms at 68 VGPRs (7 waves) and 6.7 ms at 63 (8 waves,
// with spills), against 6.5-6.6 ms here.
// The FP64 rows keep their 96 VGPRs (5 waves).
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(6))) void pf_stats_count_f32_kernel(
    const PfValueParams v, const float* __restrict__ exposures, const double* __restrict__ w,
    double* __restrict__ partials, SelRowState* __restrict__ sel, float* __restrict__ cand,
    int64_t cap, int stage, int flush, uint32_t* __restrict__ lo_parts)
{
    pf_stats_rows<float, true>(v, exposures, w, partials, sel, cand, cap, stage, flush, lo_parts);
}

// Adds each row's below-the-bracket (or zero) counts of the statistics workgroups to its
// selection state, as the counting's atomics did: one wave per row (integer sums, exact in
// any order; the row's state is written by this wave alone).
__global__ __launch_bounds__(kBlock) void sel_add_row_counts_kernel(SelRowState* __restrict__ sel,
                                                                    const uint32_t* __restrict__ parts,
                                                                    int rows, int nb)
{
    const int lane = threadIdx.x % kWave;
    const int row = blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
    if (row >= rows) return;   // whole waves
    uint32_t c = 0;
    for (int b = lane; b < nb; b += kWave) c += parts[(int64_t)row * nb + b];
    for (int o = kWave / 2; o > 0; o >>= 1) c += __shfl_xor(c, o, kWave);
    if (lane == 0 && c) {
        if (sel[row].lo)
            sel[row].below += c;
        else
            sel[row].zin += c;
    }
}

int64_t pf_stats_blocks(int64_t paths) { return (paths + kStatPaths - 1) / kStatPaths; }

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
        pf_paths_kernel<Real, false><<<grid, block, (size_t)kGroupShocks * block * sizeof(float), st>>>(
            k, g, normals, curves, logb);
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

int pf_stats_stage(int64_t paths, int64_t pfe_rank, int ndates, int elem_bytes, int* flush_dates)
{
    // per-wave staging slice: twice the expected inside count of a wave's 512 paths over
    // the dates between flushes, + 256, within 32 KB of LDS per workgroup; as many dates
    // between flushes as fit; 0 = not even one date fits
    const BracketPlan bp = bracket_plan(paths, pfe_rank, elem_bytes);
    if (!bp.use || ndates < 1) return 0;
    const double frac = (double)(bp.khi - bp.klo + 1) / bp.m;
    const double per_date = 2.0 * frac * kWave * kStatPerThread;
    const int limit = 32 * 1024 / ((kBlock / kWave) * elem_bytes);
    int f = ndates;
    if (per_date * f + 256.0 > limit) f = (int)((limit - 256.0) / per_date);
    if (f < 1) return 0;
    if (flush_dates) *flush_dates = f;
    const int want = (int)(per_date * f + 256.0);
    const int stage = (want + 63) / 64 * 64;
    return stage < limit ? stage : limit;
}

template <typename Real>
static hipError_t pf_stats_launch(const PfValueParams& v, const Real* ex, const double* w,
                                  double* partials, double* sums, void* selws, int64_t pfe_rank,
                                  uint32_t* lo_parts, hipStream_t st)
{
    const unsigned nb = (unsigned)((v.paths + kStatPaths - 1) / kStatPaths);
    const dim3 grid(nb, (unsigned)v.nsets);
    if (selws) {
        const BracketPlan bp = bracket_plan(v.paths, pfe_rank, (int)sizeof(Real));
        int flush = 0;
        const int stage = pf_stats_stage(v.paths, pfe_rank, v.ndates, (int)sizeof(Real), &flush);
        if (!bp.use || !stage) return hipErrorInvalidValue;
        SelRowState* state;
        Real* cand;
        bracket_ws_split<Real>(selws, v.ndates * v.nsets, &state, &cand);
        const size_t lds = (size_t)(kBlock / kWave) * stage * sizeof(Real);
        if constexpr (sizeof(Real) == sizeof(float))
            pf_stats_count_f32_kernel<<<grid, kBlock, lds, st>>>(v, ex, w, partials, state, cand, bp.cap,
                                                                stage, flush, lo_parts);
        else
            pf_stats_kernel<Real, true><<<grid, kBlock, lds, st>>>(v, ex, w, partials, state, cand,
                                                                  bp.cap, stage, flush, lo_parts);
        if (sizeof(Real) == sizeof(float) && lo_parts) {
            const int rows = v.ndates * v.nsets;
            const unsigned wgs = (unsigned)((rows + kBlock / kWave - 1) / (kBlock / kWave));
            sel_add_row_counts_kernel<<<wgs, kBlock, 0, st>>>(state, lo_parts, rows, (int)nb);
        }
    } else {
        pf_stats_kernel<Real, false><<<grid, kBlock, 0, st>>>(v, ex, w, partials, nullptr, nullptr, 0,
                                                             0, 0, nullptr);
    }
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return err;
    grouped_column_sum_kernel<<<dim3((unsigned)(v.ndates + 2), (unsigned)v.nsets), kBlock, 0, st>>>(
        partials, nb, v.ndates + 2, sums);
    return hipGetLastError();
}

hipError_t launch_grouped_column_sum(const double* partials, int nsets, int64_t blocks, int cols,
                                     double* sums, hipStream_t st)
{
    grouped_column_sum_kernel<<<dim3((unsigned)cols, (unsigned)nsets), kBlock, 0, st>>>(partials, blocks,
                                                                                      cols, sums);
    return hipGetLastError();
}

hipError_t launch_pf_stats_f64(const PfValueParams& v, const double* ex, const double* w,
                               double* partials, double* sums, void* selws, int64_t pfe_rank,
                               uint32_t* lo_parts, hipStream_t st)
{
    return pf_stats_launch<double>(v, ex, w, partials, sums, selws, pfe_rank, lo_parts, st);
}

hipError_t launch_pf_stats_f32(const PfValueParams& v, const float* ex, const double* w,
                               double* partials, double* sums, void* selws, int64_t pfe_rank,
                               uint32_t* lo_parts, hipStream_t st)
{
    return pf_stats_launch<float>(v, ex, w, partials, sums, selws, pfe_rank, lo_parts, st);
}

}  // namespace hjm

namespace hjm {

// ---- portfolio adjoint ----------------------------------------------------------------
// Per path J = sum_g sum_e w_ge max(V_ge, 0) e^{-lnB_e}. Seeds at each date from the
// curve checkpoint: dJ/dP_k = -sum_g w_ge 1[V_ge > 0] D_e gamma_gek (V = a - gamma.P),
// iota_k = dJ/dI_k = -P_k dJ/dP_k, S_m = sum_{k >= m} iota_k,
// gbar_l = dtau/2 (S_{l+1} [l <= 49] + S_l [l >= 1]) + [l = 50] dtau sum_{k>50} (k-50) iota_k,
// beta_e = dJ/dlnB_e = -sum_g w_ge max(V_ge, 0) D_e. Stored [date][tenor][path], [date][path].
template <typename Real>
__global__ __launch_bounds__(kBlock) void pf_adjoint_seed_kernel(const PfValueParams v,
                                                                 const Real* __restrict__ curves,
                                                                 const double* __restrict__ logb,
                                                                 const Real* __restrict__ coef,
                                                                 const Real* __restrict__ aconst,
                                                                 const int* __restrict__ kmax,
                                                                 const double* __restrict__ w,
                                                                 float* __restrict__ gbar,
                                                                 float* __restrict__ beta,
                                                                 double* __restrict__ beta_partials)
{
    const int64_t P = v.paths;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int e = blockIdx.y;
    const bool live = i < P;  // every lane takes part in the workgroup's LDS staging
    const int64_t ic = live ? i : 0;
    Real f[kTenors];
#pragma unroll
    for (int j = 0; j < kTenors; ++j) f[j] = curves[((int64_t)e * kTenors + j) * P + ic];
    const double D = exp(-logb[(int64_t)e * P + ic]);
    constexpr int KP = kMaxPeriods;
    Real Pk[KP], dP[KP];
    Real I = (Real)0;
    const Real hw = (Real)(0.5 * v.dtau);
#pragma unroll
    for (int kk = 1; kk <= KP; ++kk) {
        if (kk <= kTenors - 1) I += hw * (f[kk - 1] + f[kk]);
        else I += (Real)v.dtau * f[kTenors - 1];
        Pk[kk - 1] = exp(-I);
        dP[kk - 1] = (Real)0;
    }
    double b = 0.0;
    // one set: its value, then dP_k -= w D 1[V > 0] gamma_k
    auto set_adjoint = [&](int K, auto gam, Real V, double wg) {
        // chunks of 16 periods: each chunk's loads are issued together ahead of its FMAs
        // (one wait per 16 FMAs)
#pragma unroll
        for (int k0 = 0; k0 < KP; k0 += 16) {
            if (k0 < K) {
                Real gk[16];
#pragma unroll
                for (int u = 0; u < 16; ++u)
                    if (k0 + u < KP) gk[u] = gam[k0 + u];
#pragma unroll
                for (int u = 0; u < 16; ++u)
                    if (k0 + u < KP) V -= gk[u] * Pk[k0 + u];
            }
        }
        // dJ/dV_g = w D where the set's exposure is positive (0 elsewhere: no divergent
        // branch); dV_g/dP_k = -gamma_k
        const bool pos = V > (Real)0;
        b -= pos ? wg * (double)V * D : 0.0;
        const Real a = pos ? (Real)(wg * D) : (Real)0;
#pragma unroll
        for (int k0 = 0; k0 < KP; k0 += 16) {
            if (k0 < K) {
                Real gk[16];
#pragma unroll
                for (int u = 0; u < 16; ++u)
                    if (k0 + u < KP) gk[u] = gam[k0 + u];
#pragma unroll
                for (int u = 0; u < 16; ++u)
                    if (k0 + u < KP) dP[k0 + u] -= a * gk[u];
            }
        }
    };
    // the coefficient rows by scalar loads (FP64 plans; staging them in LDS needs more
    // registers than two waves per SIMD allow)
    kconst_ptr_pf<Real> cp = (kconst_ptr_pf<Real>)coef;
    kconst_ptr_pf<Real> ap = (kconst_ptr_pf<Real>)aconst;
    kconst_ptr_pf<int> kp = (kconst_ptr_pf<int>)kmax;
    kconst_ptr_pf<double> wp = (kconst_ptr_pf<double>)w;
    for (int gset = 0; gset < v.nsets; ++gset) {
        const int64_t row = (int64_t)e * v.nsets + gset;
        set_adjoint(kp[row], cp + row * KP, ap[row], wp[(int64_t)gset * v.ndates + e]);
    }
    // suffix sums of iota_k = -P_k dP_k and the extrapolation moment
    Real S = (Real)0, ext = (Real)0;
    Real Sk[KP + 2];
    Sk[KP + 1] = (Real)0;
    Sk[KP] = (Real)0;
#pragma unroll
    for (int kk = KP; kk >= 1; --kk) {
        const Real iota = -Pk[kk - 1] * dP[kk - 1];
        S += iota;
        Sk[kk] = S;
        if (kk > kTenors - 1) ext += (Real)(kk - (kTenors - 1)) * iota;
    }
    // the workgroup's sum of beta at this date (the drift gradient's bank-account terms)
    {
        __shared__ double bred[kBlock / kWave];
        const double bs = wave_sum(live ? (double)(float)b : 0.0);
        if (threadIdx.x % kWave == 0) bred[threadIdx.x / kWave] = bs;
        __syncthreads();
        if (threadIdx.x == 0) {
            double t = 0.0;
#pragma unroll
            for (int q = 0; q < kBlock / kWave; ++q) t += bred[q];
            beta_partials[(int64_t)blockIdx.x * v.ndates + e] = t;
        }
    }
    if (!live) return;
    float* gb = gbar + (int64_t)e * kTenors * P + i;
#pragma unroll
    for (int l = 0; l < kTenors; ++l) {
        Real val = (Real)0;
        if (l <= kTenors - 2) val += hw * Sk[l + 1];
        if (l >= 1) val += hw * Sk[l];
        if (l == kTenors - 1) val += (Real)v.dtau * ext;
        gb[(int64_t)l * P] = (float)val;
    }
    beta[(int64_t)e * P + i] = (float)b;
}

// Reverse sweep over the whole horizon (n = S-1 .. 0) in one pass, two lanes per path
// (the lower half-wave holds tenors 0..25, the upper half-wave tenors 25..50 with tenor 25
// copied from the lower lane each step): loading gradients with lambda(n+1) and the shocks
// of step n, lambda <- T^T lambda + mu e_0 (mu = dt sum of beta of later dates), + gbar_e at
// the date steps. The drift gradient comes from the block-end adjoints lambda_e, summed
// over dates and paths (the host applies sum_{k<S} (T^T)^k to them and the mu terms).
// kSrc: 0 regenerate the in-kernel normals, 1 host-supplied FP64 [path][step][3],
// 2 the forward pass's stored FP32 normals [step][3][path].
// partials[workgroup][5 * 51]: g0, g1, g2, sum of lambda_e, d J / d spot.
template <int kSrc>
__global__ __launch_bounds__(kBlock) void pf_adjoint_kernel(const PfGrid g, float c,
                                                            const double* __restrict__ normals,
                                                            const float* __restrict__ zs,
                                                            const float* __restrict__ gbar,
                                                            const float* __restrict__ beta,
                                                            double* __restrict__ partials)
{
    __shared__ double acc[kBlock / kWave][2][kAdjKinds][kHalfSlots];
    __shared__ float lacc[kHalfSlots][kBlock];  // per lane: sum over dates of lambda_e
    __shared__ float zbuf[kSrc == 0 ? kGroupShocks * kAdjPaths : 1];
    const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
    const int half = lane >> 5;
    const int pslot = wave * 32 + (lane & 31);
    for (int k = threadIdx.x; k < (kBlock / kWave) * 2 * kAdjKinds * kHalfSlots; k += kBlock)
        (&acc[0][0][0][0])[k] = 0.0;
#pragma unroll
    for (int k = 0; k < kHalfSlots; ++k) lacc[k][threadIdx.x] = 0.f;
    __syncthreads();
    const int64_t P = g.paths;
    const int64_t i = (int64_t)blockIdx.x * kAdjPaths + pslot;
    const bool active = i < P;
    const int64_t ic = active ? i : 0;
    const uint32_t pid = (uint32_t)(g.path0 + ic);
    float* col = zbuf + (kSrc == 0 ? pslot : 0);
    const bool upper = half == 1;
    const float sqdt = (float)sqrt(g.dt), dtf = (float)g.dt;
    float cv = c, omc = 1.0f - c, a25 = upper ? 1.0f + c : 1.0f - c, b24 = upper ? -c : 0.f;
    asm volatile("" : "+v"(cv), "+v"(omc), "+v"(a25), "+v"(b24));
    const int tbase = 25 * half;
    // a lane adds the seeds of its own tenors (the upper lane's copy of tenor 25 takes none)
    const bool seeded = active;
    auto add_seeds = [&](int e, float (&l)[kHalfSlots]) {
        const float* q = gbar + ((int64_t)e * kTenors + tbase) * P + ic;
#pragma unroll
        for (int k = 0; k < kHalfSlots; ++k)
            l[k] += (seeded && !(upper && k == 0)) ? q[(int64_t)k * P] : 0.f;
    };
    float lam[kHalfSlots], g0[kHalfSlots], g1[kHalfSlots], g2[kHalfSlots];
#pragma unroll
    for (int k = 0; k < kHalfSlots; ++k) lam[k] = g0[k] = g1[k] = g2[k] = 0.f;
    const int E = g.ndates_m1;
    add_seeds(E, lam);
    float mu = 0.f;

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
    auto step = [&](float z0, float z1, float z2) {
        const float s0 = sqdt * z0, s1 = sqdt * z1, s2 = sqdt * z2;
#pragma unroll
        for (int k = 0; k < kHalfSlots; ++k) {
            op_fmac_vv(g0[k], lam[k], s0);
            op_fmac_vv(g1[k], lam[k], s1);
            op_fmac_vv(g2[k], lam[k], s2);
        }
        const uint32_t l25 = __builtin_bit_cast(uint32_t, lam[kHalfSlots - 1]);
        const float recv = __builtin_bit_cast(float, __builtin_amdgcn_permlane32_swap(l25, l25, false, false)[0]);
        lam[0] = upper ? recv : lam[0];
        // lambda <- T^T lambda (+ mu on tenor 0), descending
        float t25, t24;
        op_mul_vv(t25, cv, lam[24]);
        op_mul_vv(t24, cv, lam[23]);
        op_fma_vvv(t24, b24, lam[25], t24);
        op_fma_inplace(lam[25], a25, t25);
        op_fma_inplace(lam[24], omc, t24);
        constexpr int G = 4;
#pragma unroll
        for (int k0 = 23; k0 >= 1; k0 -= G) {
            float t[G];
#pragma unroll
            for (int u = 0; u < G; ++u)
                if (k0 - u >= 1) op_mul_vv(t[u], cv, lam[k0 - u - 1]);
#pragma unroll
            for (int u = 0; u < G; ++u)
                if (k0 - u >= 1) op_fma_inplace(lam[k0 - u], omc, t[u]);
        }
        op_fma_inplace(lam[0], omc, mu);
    };

    const int S = g.stride;
    constexpr int kPre = 5;
    float zq[kPre][3];
    auto loadz = [&](int n, float (&d)[3]) {
        const float* q = zs + (int64_t)n * 3 * P + ic;
        d[0] = q[0];
        d[1] = q[P];
        d[2] = q[2 * P];
    };
    if constexpr (kSrc == 2) {
#pragma unroll
        for (int u = 0; u < kPre; ++u) loadz(E * S - 1 - u, zq[u]);
    }
    const int whole = S / kGroupSteps * kGroupSteps;
    for (int e = E - 1; e >= 0; --e) {
        mu += (active && !upper) ? dtf * beta[(int64_t)(e + 1) * P + ic] : 0.f;
#pragma unroll
        for (int k = 0; k < kHalfSlots; ++k) lacc[k][threadIdx.x] += lam[k];
        if constexpr (kSrc == 2) {
            for (int s0 = S - 1; s0 >= 0; s0 -= kPre) {
#pragma unroll
                for (int u = 0; u < kPre; ++u) {
                    const float a0 = zq[u][0], a1 = zq[u][1], a2 = zq[u][2];
                    const int nn = e * S + s0 - kPre - u;
                    if (nn >= 0) loadz(nn, zq[u]);
                    step(a0, a1, a2);
                }
            }
        } else {
            for (int s = S - 1; s >= 0; --s) {
                const int n = e * S + s;
                if constexpr (kSrc == 1) {
                    const double* q = normals + (ic * g.nsteps + n) * 3;
                    step((float)q[0], (float)q[1], (float)q[2]);
                } else {
                    Normals3 z;
                    if (s < whole) {
                        if ((s & (kGroupSteps - 1)) == kGroupSteps - 1) {
                            __syncthreads();
                            fill_group_half(g.seed_lo, g.seed_hi, pid, (uint32_t)e, s / kGroupSteps,
                                            half, col, kAdjPaths);
                            __syncthreads();
                        }
                        z = group_step(col, kAdjPaths, s);
                    } else {
                        z = step_normals(g.seed_lo, g.seed_hi, pid, (uint32_t)n, (uint32_t)S);
                    }
                    step(z.z0, z.z1, z.z2);
                }
            }
        }
        // date e: the valuation at e depends on f(n_e)
        add_seeds(e, lam);
        if ((e & 1) == 0) {
            flush(g0, 0);
            flush(g1, 1);
            flush(g2, 2);
        }
    }
    {
        float la[kHalfSlots];
#pragma unroll
        for (int k = 0; k < kHalfSlots; ++k) la[k] = lacc[k][threadIdx.x];
        flush(la, 3);
        flush(lam, 4);
    }
    __syncthreads();
    for (int o = threadIdx.x; o < kAdjKinds * kTenors; o += kBlock) {
        const int kind = o / kTenors, t = o % kTenors;
        const int h = t > 25 ? 1 : 0, k = t - 25 * h;
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < kBlock / kWave; ++w) s += acc[w][h][kind][k];
        partials[(int64_t)blockIdx.x * (kAdjKinds * kTenors) + o] = s;
    }
}

template <typename Real>
static hipError_t pf_adj_seed_launch(const PfValueParams& v, const Real* curves, const double* logb,
                                     const Real* coef, const Real* a, const int* kmax,
                                     const double* w, float* gbar, float* beta,
                                     double* beta_partials, hipStream_t st)
{
    const dim3 grid((unsigned)((v.paths + kBlock - 1) / kBlock), (unsigned)v.ndates);
    pf_adjoint_seed_kernel<Real><<<grid, kBlock, 0, st>>>(v, curves, logb, coef, a, kmax, w, gbar,
                                                          beta, beta_partials);
    return hipGetLastError();
}

hipError_t launch_pf_adjoint_seed_f64(const PfValueParams& v, const double* curves,
                                      const double* logb, const double* coef, const double* a,
                                      const int* kmax, const double* w, float* gbar, float* beta,
                                      double* beta_partials, hipStream_t st)
{
    return pf_adj_seed_launch<double>(v, curves, logb, coef, a, kmax, w, gbar, beta, beta_partials,
                                      st);
}

int64_t pf_adjoint_blocks(int64_t paths) { return (paths + kAdjPaths - 1) / kAdjPaths; }

hipError_t launch_pf_adjoint(const PfGrid& g, double c, const double* normals, const float* zs,
                             const float* gbar, const float* beta, double* partials, hipStream_t st)
{
    const dim3 grid((unsigned)pf_adjoint_blocks(g.paths));
    const float cf = step_ratio_f32(c);  // 1 - cf, 1 + cf exact, as in the FP32 forward
    if (zs && g.stride % 5 != 0) return hipErrorInvalidValue;  // stored-normal read-ahead
    if (normals)
        pf_adjoint_kernel<1><<<grid, kBlock, 0, st>>>(g, cf, normals, zs, gbar, beta, partials);
    else if (zs)
        pf_adjoint_kernel<2><<<grid, kBlock, 0, st>>>(g, cf, normals, zs, gbar, beta, partials);
    else
        pf_adjoint_kernel<0><<<grid, kBlock, 0, st>>>(g, cf, normals, zs, gbar, beta, partials);
    return hipGetLastError();
}

}  // namespace hjm
