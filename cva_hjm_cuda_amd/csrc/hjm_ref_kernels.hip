// hjm_ref_kernels.hip -- reference-semantics HJM path generation and exposure, gfx950.
//
// Layout: one Monte Carlo path per lane (wave64, 256-lane workgroups). The 51-point
// forward curve and its 51 discount integrals live in VGPRs for the whole horizon; the
// tenor constants (drift*dt, 3 loadings) come from the kernarg segment through scalar
// loads, identical for every lane. Normals are generated in-kernel (counter Philox) or
// read from the caller's [path][step][3] array. HBM is touched only at exposure dates:
// one coalesced store per numeraire slot, slot-major [slot][path].
#include <algorithm>

#include "hjm_common.h"
#include "hjm_rng.h"
#include "hjm_select.h"
#include "hjm_kernels.h"
#include "hjm_valu.h"

namespace hjm {

template <typename Real>
__device__ __forceinline__ Real fma_r(Real a, Real b, Real c)
{
    return __builtin_fma(a, b, c);
}
template <>
__device__ __forceinline__ float fma_r<float>(float a, float b, float c)
{
    return __builtin_fmaf(a, b, c);
}

template <typename T>
using kconst_ptr = const __attribute__((address_space(4))) T*;

// The step constants are the kernel's first argument, i.e. they start at offset 0 of
// the kernarg segment. Reading them through this pointer, re-laundered every Euler step,
// keeps the compiler from hoisting all 204 of them out of the step loop (they do not
// fit in SGPRs and would be spilled to VGPR lanes): each step re-reads them with scalar
// loads, which cost no vector-ALU issue slots.
template <typename Real>
__device__ __forceinline__ kconst_ptr<RefStepConsts<Real>> step_consts_ptr()
{
    return (kconst_ptr<RefStepConsts<Real>>)__builtin_amdgcn_kernarg_segment_ptr();
}

// Path generation of simulation_cpu.h:188-262 with the driver loop :409-436
// (equivalently simulation_gpu.cu:471-669): block 0 (steps 0..stride-1) keeps the
// initial curve; blocks 1..ndates-1 take Euler steps
//   f_j <- f_j + m_j dt + sqrt(dt) sum_k sigma_kj phi_k + (f_{j+1} - f_j) dt / dtau
// (backward difference at j = 50), every step adds f_j into the integral A_j, and
// the end of block b records slot b+1: x = f_{b+1}, y = exp(-A_{b+1} dt).
//
// kCompensated (FP32): the per-step sums go into an FP32 block partial P_j (at most
// `stride` terms), which is Kahan-added into (A_j, C_j) at the end of every block; the
// recorded integral is A_j - C_j in FP64. Otherwise A_j is a plain FP64 running sum.
//
// Per tenor and step the update is five FMAs and one add, ordered so that the old
// neighbour f_{j+1} is read first and the last FMA rewrites f_j in place: no copies.
template <typename Real, bool kCompensated, bool kHostNormals>
__global__ __launch_bounds__(kBlock, 2) void ref_paths_kernel(const RefStepConsts<Real> k,
                                                           const RefGrid g,
                                                           const double* __restrict__ normals,
                                                           double* __restrict__ num_x,
                                                           double* __restrict__ num_y)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= g.paths) return;
    const int64_t P = g.paths;
    const uint32_t pid = (uint32_t)(g.path0 + i);
    const Real sqdt = k.sqrt_dt;
    // c, -c, 1-c, 1+c in VGPRs: each update FMA may then take its SGPR tenor constant
    // (one scalar operand per VALU instruction).
    Real cv = k.c, mcv = -k.c, omc = k.omc, opc = k.opc;
    asm volatile("" : "+v"(cv), "+v"(mcv), "+v"(omc), "+v"(opc));

    Real f[kTenors], A[kTenors], C[kTenors], Pb[kTenors];
#pragma unroll
    for (int j = 0; j < kTenors; ++j) {
        f[j] = k.spot[j];
        A[j] = (Real)g.a0[j];
        C[j] = kCompensated ? (Real)((double)A[j] - g.a0[j]) : (Real)0;
        Pb[j] = (Real)0;
    }
    num_x[P + i] = (double)k.spot[1];
    num_y[P + i] = exp(-g.a0[1] * g.dt);

    // each lane's column of staged shocks (16 steps at a time)
    __shared__ float zbuf[kHostNormals ? 1 : kGroupShocks * kBlock];
    float* col = zbuf + (kHostNormals ? 0 : threadIdx.x);
    const int whole = g.stride / kGroupSteps * kGroupSteps;
    for (int b = 1; b < g.ndates; ++b) {
        for (int s = 0; s < g.stride; ++s) {
            const int n = b * g.stride + s;
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
                        fill_group(g.seed_lo, g.seed_hi, pid, (uint32_t)b, s / kGroupSteps, col, kBlock);
                    z = group_step(col, kBlock, s);
                } else {
                    z = step_normals(g.seed_lo, g.seed_hi, pid, (uint32_t)n, (uint32_t)g.stride);
                }
                z0 = (Real)z.z0;
                z1 = (Real)z.z1;
                z2 = (Real)z.z2;
            }
            const Real s0 = sqdt * z0, s1 = sqdt * z1, s2 = sqdt * z2;
            kconst_ptr<RefStepConsts<Real>> kp = step_consts_ptr<Real>();
            asm volatile("" : "+s"(kp));
            // last tenor: backward difference, its neighbour term taken while f_49 is old
            constexpr int L = kTenors - 1;
            Real t50;
            op_fma_vvs(t50, mcv, f[L - 1], kp->ddt[L]);
            // tenors 0..49 in groups of 4: each instruction of the sequence for 4 tenors
            // before the next, so consecutive instructions are independent
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
#pragma unroll
                for (int u = 0; u < G; ++u)
                    if (j0 + u < L) op_add_inplace(kCompensated ? Pb[j0 + u] : A[j0 + u], f[j0 + u]);
            }
            op_fmac_sv(t50, kp->v0[L], s0);
            op_fmac_sv(t50, kp->v1[L], s1);
            op_fmac_sv(t50, kp->v2[L], s2);
            op_fma_inplace(f[L], opc, t50);
            op_add_inplace(kCompensated ? Pb[L] : A[L], f[L]);
        }
        if constexpr (kCompensated) {
#pragma unroll
            for (int j = 0; j < kTenors; ++j) {
                const Real y = Pb[j] - C[j];
                const Real t = A[j] + y;
                C[j] = (t - A[j]) - y;
                A[j] = t;
                Pb[j] = (Real)0;
            }
        }
        const int slot = b + 1;
        Real x = f[0], a = A[0], cc = C[0];
#pragma unroll
        for (int j = 1; j < kTenors; ++j) {
            if (j == slot) {
                x = f[j];
                a = A[j];
                cc = C[j];
            }
        }
        const double integral = kCompensated ? (double)a - (double)cc : (double)a;
        num_x[(int64_t)slot * P + i] = (double)x;
        num_y[(int64_t)slot * P + i] = exp(-integral * g.dt);
    }
}

// Exposure of simulation_cpu.h:269-296 (simulation_gpu.cu:803-858) for one path per
// lane: CF_i = DF_i N acc_i (m (e^{x_i/m} - 1) - K), E_t = max(sum_{i>t} CF_i, 0),
// plus the per-path CVA sample sum_t w_t E_t. Writes E [date][path] when asked and one
// row of deterministic workgroup partial sums (EE numerators, CVA, CVA^2).
__global__ __launch_bounds__(kBlock) void ref_exposure_kernel(const ExposureParams e,
                                                              const double* __restrict__ num_x,
                                                              const double* __restrict__ num_y,
                                                              double* __restrict__ exposures,
                                                              double* __restrict__ partials)
{
    __shared__ double red[kBlock / kWave][kStatSlots];
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const bool active = i < e.paths;
    const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
    const int64_t P = e.paths;

    // dates t >= ndates hold no cash flow: E_t = 0 (their exposure rows are zeroed once
    // when the plan is created)
    double V = 0.0, cva = 0.0;
    for (int t = e.ndates - 1; t >= 0; --t) {
        double Et = 0.0;
        if (active) {
            const int slot = t + 1;
            const double x = num_x[(int64_t)slot * P + i];
            const double y = num_y[(int64_t)slot * P + i];
            const double libor = e.m * (exp(x / e.m) - 1.0);
            const double cf = y * e.notional * e.accrual[slot] * (libor - e.K);
            V += cf;
            Et = V > 0.0 ? V : 0.0;
        }
        if (active && exposures) exposures[(int64_t)t * P + i] = Et;
        cva += e.w[t] * Et;
        const double ws = wave_sum(Et);
        if (lane == 0) red[wave][t] = ws;
    }
    if (lane == 0)
        for (int t = e.ndates; t < kTenors; ++t) red[wave][t] = 0.0;
    const double c1 = wave_sum(cva), c2 = wave_sum(cva * cva);
    if (lane == 0) {
        red[wave][kTenors] = c1;
        red[wave][kTenors + 1] = c2;
    }
    __syncthreads();
    if (threadIdx.x < kStatSlots) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < kBlock / kWave; ++w) s += red[w][threadIdx.x];
        partials[(int64_t)blockIdx.x * kStatSlots + threadIdx.x] = s;
    }
}

// Column sums of a [rows][cols] FP64 array in a fixed order: workgroup c sums column c.
__global__ __launch_bounds__(kBlock) void column_sum_kernel(const double* __restrict__ partials,
                                                            int64_t rows, int cols,
                                                            double* __restrict__ out)
{
    __shared__ double red[kBlock / kWave];
    const int c = blockIdx.x;
    double s = 0.0;
    for (int64_t r = threadIdx.x; r < rows; r += kBlock) s += partials[r * cols + c];
    s = wave_sum(s);
    if (threadIdx.x % kWave == 0) red[threadIdx.x / kWave] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
#pragma unroll
        for (int w = 0; w < kBlock / kWave; ++w) t += red[w];
        out[c] = t;
    }
}

// The in-kernel generator's normals for [path0, path0 + paths) x [0, nsteps), [path][step][3].
// A workgroup takes np <= 64 paths of one date block (lane l < np of every wave is path l;
// np = 64 unless the paths' LDS rows would exceed 64 KB):
//   * the block's whole 16-step groups: 12 calls per path and group (calls 3 m .. 3 m + 2
//     of the four classes), the wave-uniform work items (group, class, call) spread over
//     the four waves;
//   * the steps past the last whole group (2 of 50): one work item per step, computed
//     directly;
// each into the path's LDS row; then every path's stride x 3 values go out as one
// contiguous run (coalesced across the lanes).
__global__ __launch_bounds__(kBlock) void normals_kernel(uint32_t seed_lo, uint32_t seed_hi,
                                                         int64_t path0, int64_t paths,
                                                         int nsteps, int stride, int np,
                                                         double* __restrict__ out)
{
    extern __shared__ float nbuf[];  // [np][stride * 3]
    const int nblk = (nsteps + stride - 1) / stride;
    const int b = (int)(blockIdx.x % nblk);
    const int64_t pbase = (int64_t)(blockIdx.x / nblk) * np;
    const int lane = threadIdx.x % kWave, wv = threadIdx.x / kWave, nwv = blockDim.x / kWave;
    const int row = 3 * stride;
    const int64_t p = pbase + lane;
    const uint32_t pid = (uint32_t)(path0 + (p < paths ? p : 0));
    // lanes past np compute like the others but store nothing
    float* mine = nbuf + (lane < np ? lane : 0) * row;
    const bool owner = lane < np;
    const int ngroups = stride / kGroupSteps, whole = ngroups * kGroupSteps;
    // whole groups: item = (group m, class g, call cc), uniform across the wave
    for (int it = wv; it < ngroups * 12; it += nwv) {
        const int m = it / 12, g = (it / 3) % 4, cc = it % 3;
        float z[4];
        stream_call(seed_lo, seed_hi, pid, (uint32_t)b, (uint32_t)g, (uint32_t)(3 * m + cc), z);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int jj = 4 * cc + e;  // shock of class g within the group
            const int s = kGroupSteps * m + 4 * (jj / 3) + g;
            if (owner) mine[3 * s + jj % 3] = z[e];
        }
    }
    // the steps past the last whole group
    for (int s = whole + wv; s < stride; s += nwv) {
        const Normals3 z = step_normals(seed_lo, seed_hi, pid, (uint32_t)(b * stride + s), (uint32_t)stride);
        if (owner) {
            mine[3 * s + 0] = z.z0;
            mine[3 * s + 1] = z.z1;
            mine[3 * s + 2] = z.z2;
        }
    }
    __syncthreads();
    const int len = 3 * min(stride, nsteps - b * stride);
    for (int l = 0; l < np; ++l) {
        if (pbase + l >= paths) break;
        double* o = out + ((pbase + l) * nsteps + (int64_t)b * stride) * 3;
        for (int r = threadIdx.x; r < len; r += blockDim.x) o[r] = (double)nbuf[l * row + r];
    }
}

// ---- launchers -------------------------------------------------------------------
template <typename Real, bool kKahan>
static hipError_t launch_paths(const RefStepConsts<Real>& k, const RefGrid& g,
                               const double* normals, double* nx, double* ny, hipStream_t st)
{
    const dim3 grid((unsigned)((g.paths + kBlock - 1) / kBlock));
    if (normals)
        ref_paths_kernel<Real, kKahan, true><<<grid, kBlock, 0, st>>>(k, g, normals, nx, ny);
    else
        ref_paths_kernel<Real, kKahan, false><<<grid, kBlock, 0, st>>>(k, g, normals, nx, ny);
    return hipGetLastError();
}

hipError_t launch_ref_paths_f64(const RefStepConsts<double>& k, const RefGrid& g,
                                const double* normals, double* nx, double* ny, hipStream_t st)
{
    return launch_paths<double, false>(k, g, normals, nx, ny, st);
}

hipError_t launch_ref_paths_f32k(const RefStepConsts<float>& k, const RefGrid& g,
                                 const double* normals, double* nx, double* ny, hipStream_t st)
{
    return launch_paths<float, true>(k, g, normals, nx, ny, st);
}

hipError_t launch_ref_exposure(const ExposureParams& e, const double* nx, const double* ny,
                               double* exposures, double* partials, hipStream_t st)
{
    const dim3 grid((unsigned)((e.paths + kBlock - 1) / kBlock));
    ref_exposure_kernel<<<grid, kBlock, 0, st>>>(e, nx, ny, exposures, partials);
    return hipGetLastError();
}

hipError_t launch_column_sum(const double* partials, int64_t rows, int cols, double* out,
                             hipStream_t st)
{
    column_sum_kernel<<<cols, kBlock, 0, st>>>(partials, rows, cols, out);
    return hipGetLastError();
}

// The Box-Muller factors of n words: r[i] = radius of words[i] as the first word of a pair,
// c[i], s[i] = cosine and sine of words[i] as the second.
__global__ __launch_bounds__(kBlock) void box_muller_parts_kernel(const uint32_t* __restrict__ words,
                                                                  int64_t n, float* __restrict__ r,
                                                                  float* __restrict__ c,
                                                                  float* __restrict__ s)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const uint32_t w = words[i];
    r[i] = bm_radius(w);
    bm_angle(w, c[i], s[i]);
}

hipError_t launch_box_muller_parts(const uint32_t* words, int64_t n, float* r, float* c, float* s,
                                   hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    box_muller_parts_kernel<<<(unsigned)((n + kBlock - 1) / kBlock), kBlock, 0, st>>>(words, n, r, c, s);
    return hipGetLastError();
}

hipError_t launch_normals(uint64_t seed, int64_t path0, int64_t paths, int nsteps, int stride,
                          double* out, hipStream_t st)
{
    // LDS: np paths x stride x 3 floats, within 64 KB
    if (stride <= 0 || 12 * (int64_t)stride > 65536) return hipErrorInvalidValue;
    const int np = (int)std::min<int64_t>(kWave, 65536 / (12 * (int64_t)stride));
    const int64_t nblk = (nsteps + stride - 1) / stride;
    const int64_t groups = (paths + np - 1) / np;
    const dim3 grid((unsigned)(groups * nblk));
    normals_kernel<<<grid, kBlock, (size_t)np * 12 * stride, st>>>(
        (uint32_t)seed, (uint32_t)(seed >> 32), path0, paths, nsteps, stride, np, out);
    return hipGetLastError();
}

}  // namespace hjm
