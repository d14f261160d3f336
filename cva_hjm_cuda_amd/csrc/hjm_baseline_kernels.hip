// hjm_baseline_kernels.hip -- the reference's GPU design on gfx950, kept as a baseline.
//
// This is the "plain FP64 GPU form" that the optimised path kernel is measured against:
// the decomposition of simulation_gpu.cu (normals materialised in HBM by a separate
// generator pass, :193-216; one workgroup per path with one tenor per thread and the
// curve in LDS, 50 sub-steps and two barriers per sub-step per launch, state and
// integrals round-tripping through HBM between launches, :471-669; exposure with a
// serial suffix sum per thread, :803-858; a column-sum pass for EE, :889-933) written
// as plain HIP with FP64 arithmetic in the oracle's operation order. It is correct (its
// outputs match the oracle) but is not meant to be fast.
#include "hjm_common.h"
#include "hjm_kernels.h"

namespace hjm {

constexpr int kTenorBlock = 64;  // one wave per path, lanes 51..63 idle

// One 50-step block of one path per workgroup (simulation_gpu.cu:471-587 structure,
// simulation_cpu.h:37-54 arithmetic, no contraction).
__global__ __launch_bounds__(kTenorBlock) void blockwise_paths_kernel(
    const RefStepConsts<double> k, const RefGrid g, int b, const double* __restrict__ normals,
    const double* __restrict__ state_in, double* __restrict__ state_out,
    double* __restrict__ integral, double* __restrict__ nx, double* __restrict__ ny)
{
#pragma clang fp contract(off)
    __shared__ double curve[kTenorBlock];
    const int t = threadIdx.x;
    const int64_t p = blockIdx.x;
    const bool on = t < kTenors;
    const int64_t gi = p * kTenors + t;
    double sum = on ? integral[gi] : 0.0;
    if (on) curve[t] = (b > 0) ? state_in[gi] : k.spot[t];
    __syncthreads();
    const double dt = g.dt;
    const double dtau = k.dtau;
    const double sq = sqrt(dt);
    double rate = 0.0;
    for (int s = 0; s < g.stride; ++s) {
        if (on) {
            if (b == 0) {
                rate = k.spot[t];
            } else {
                const double dF = (t == kTenors - 1) ? curve[t] - curve[t - 1] : curve[t + 1] - curve[t];
                const double* phi = normals + (p * g.nsteps + (int64_t)b * g.stride + s) * 3;
                double vol_sum = k.v0[t] * phi[0];
                vol_sum += k.v1[t] * phi[1];
                vol_sum += k.v2[t] * phi[2];
                vol_sum *= sq;
                double dfbar = k.drift[t] * dt;
                dfbar += vol_sum;
                dfbar += (dF / dtau) * dt;
                rate = curve[t] + dfbar;
            }
            sum += rate;
        }
        __syncthreads();
        if (on) curve[t] = rate;
        __syncthreads();
    }
    if (on) {
        state_out[gi] = rate;
        integral[gi] = sum;
        if (t == b + 1) {
            const int64_t P = g.paths;
            nx[(int64_t)t * P + p] = rate;
            ny[(int64_t)t * P + p] = exp(-sum * dt);
        }
    }
}

// Exposure per path per workgroup, serial suffix sums (simulation_gpu.cu:803-858),
// output path-major [path][51] as the reference stores it.
__global__ __launch_bounds__(kTenorBlock) void blockwise_exposure_kernel(
    const ExposureParams e, const double* __restrict__ nx, const double* __restrict__ ny,
    double* __restrict__ exposures)
{
#pragma clang fp contract(off)
    __shared__ double cf[kTenors];
    const int t = threadIdx.x;
    const int64_t p = blockIdx.x;
    const int64_t P = e.paths;
    if (t < kTenors) {
        double c = 0.0;
        if (t >= 1 && t <= e.ndates) {
            const double x = nx[(int64_t)t * P + p];
            const double y = ny[(int64_t)t * P + p];
            const double libor = e.m * (exp(x / e.m) - 1.0);
            c = y * e.notional * e.accrual[t] * (libor - e.K);
        }
        cf[t] = c;
    }
    __syncthreads();
    if (t < kTenors) {
        double sum = 0.0;
        for (int i = t + 1; i < kTenors; ++i) sum += cf[i];
        exposures[p * kTenors + t] = sum > 0.0 ? sum : 0.0;
    }
}

// EE numerators: column sums of the [paths][51] exposure matrix (the reference's
// cublasDgemv with a ones vector, simulation_gpu.cu:912-933). Workgroup per column.
__global__ __launch_bounds__(kBlock) void column_mean_gemv_kernel(const double* __restrict__ ex,
                                                                  int64_t paths,
                                                                  double* __restrict__ sums)
{
    __shared__ double red[kBlock / kWave];
    const int c = blockIdx.x;
    double s = 0.0;
    for (int64_t r = threadIdx.x; r < paths; r += kBlock) s += ex[r * kTenors + c];
    s = wave_sum(s);
    if (threadIdx.x % kWave == 0) red[threadIdx.x / kWave] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < kBlock / kWave; ++w) t += red[w];
        sums[c] = t;
    }
}

hipError_t launch_blockwise_paths(const RefStepConsts<double>& k, const RefGrid& g, int b,
                                  const double* normals, const double* state_in, double* state_out,
                                  double* integral, double* nx, double* ny, hipStream_t st)
{
    blockwise_paths_kernel<<<(unsigned)g.paths, kTenorBlock, 0, st>>>(k, g, b, normals, state_in,
                                                                       state_out, integral, nx, ny);
    return hipGetLastError();
}

hipError_t launch_blockwise_exposure(const ExposureParams& e, const double* nx, const double* ny,
                                     double* exposures, hipStream_t st)
{
    blockwise_exposure_kernel<<<(unsigned)e.paths, kTenorBlock, 0, st>>>(e, nx, ny, exposures);
    return hipGetLastError();
}

hipError_t launch_column_mean_gemv(const double* exposures, int64_t paths, double* sums,
                                   hipStream_t st)
{
    column_mean_gemv_kernel<<<kTenors, kBlock, 0, st>>>(exposures, paths, sums);
    return hipGetLastError();
}

}  // namespace hjm
