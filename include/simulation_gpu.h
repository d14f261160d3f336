/*
 * simulation_gpu.h -- drop-in for the reference's GPU entry points
 * (simulation_gpu.h:10,12 of mahsanchez/cva_hjm_cuda), implemented on the C ABI of
 * libhjm_cva.so (hjm_cva.h). Link with -lhjm_cva.
 *
 * Same names, signatures and argument meaning: the 51-point expected-exposure profile of
 * a payer IRS under the HJM Musiela SDE is written to expected_exposure; the scenarios
 * are split over num_gpus contiguous shards; `scale` is unused. Differences, all
 * documented in the library header: normals come from the in-kernel Philox generator
 * keyed by seed 1234 instead of cuRAND; the remainder scenarios % num_gpus is simulated
 * instead of dropped; on failure the error is printed to stderr (as the reference does)
 * and the output is filled with NaN rather than left as whatever the device produced.
 * The status code of the last call on this thread is hjm_dropin::last_status().
 */
#ifndef HJM_CVA_SIMULATION_GPU_H
#define HJM_CVA_SIMULATION_GPU_H

#include <cmath>
#include <cstdio>

#include "hjm_cva.h"
#include "product.h"

namespace hjm_dropin {
inline int& last_status()
{
    static thread_local int s = 0;
    return s;
}
template <typename real>
inline void report(int status, real* out)
{
    last_status() = status;
    if (status != HJM_OK) {
        std::fprintf(stderr, "ERROR: calculateExposureMultiGPU failed: %s (%d)\n",
                     hjm_status_string(status), status);
        for (int t = 0; t < HJM_TENORS; ++t) out[t] = (real)NAN;
    }
}
}  // namespace hjm_dropin

inline void calculateExposureMultiGPU(double* expected_exposure, InterestRateSwap<double> payOff,
                                      double* accrual, double* spot_rates, double* drift,
                                      double* volatilities, double scale, int num_gpus, int simN,
                                      double dt)
{
    const int st = hjm_calculate_exposure_multi_gpu_f64(
        expected_exposure, payOff.notional, payOff.K, payOff.expiry, payOff.dtau, accrual,
        spot_rates, drift, volatilities, scale, num_gpus, simN, dt);
    if (expected_exposure) hjm_dropin::report(st, expected_exposure);
}

inline void calculateExposureMultiGPU(float* expected_exposure, InterestRateSwap<float> payOff,
                                      float* accrual, float* spot_rates, float* drifts,
                                      float* volatilities, float scale, const int num_gpus,
                                      int scenarios, float dt)
{
    const int st = hjm_calculate_exposure_multi_gpu_f32(
        expected_exposure, payOff.notional, payOff.K, payOff.expiry, payOff.dtau, accrual,
        spot_rates, drifts, volatilities, scale, num_gpus, scenarios, dt);
    if (expected_exposure) hjm_dropin::report(st, expected_exposure);
}

#endif
