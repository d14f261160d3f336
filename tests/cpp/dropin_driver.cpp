// A caller written against the reference's API (simulation_gpu.h / product.h), built
// against this library's drop-in headers. Reads the 51-point market arrays from stdin
// (spot, drift, 153 vols, accrual), runs calculateExposureMultiGPU in double and float
// like the reference's driver (simulation_gpu_testing.cpp:200-252) and prints both EE
// profiles, one value per line with %.17g. Args: scenarios num_gpus expiry.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "simulation_gpu.h"

int main(int argc, char** argv)
{
    const int scenarios = argc > 1 ? std::atoi(argv[1]) : 5000;
    const int num_gpus = argc > 2 ? std::atoi(argv[2]) : 1;
    const double expiry = argc > 3 ? std::atof(argv[3]) : 25.0;
    std::vector<double> spot(51), drift(51), vols(153), acc(51);
    for (auto* v : {&spot, &drift, &vols, &acc})
        for (double& x : *v)
            if (std::scanf("%lf", &x) != 1) return 3;
    std::vector<double> sched(51);
    for (int i = 0; i < 51; ++i) sched[i] = 0.5 * i;
    InterestRateSwap<double> payOff(sched.data(), sched.data(), sched.data(), 10.0, 0.067, expiry,
                                    0.5);
    std::vector<double> ee(51);
    calculateExposureMultiGPU(ee.data(), payOff, acc.data(), spot.data(), drift.data(), vols.data(),
                              0.0, num_gpus, scenarios, 0.01);
    if (hjm_dropin::last_status() != HJM_OK) return 1;

    std::vector<float> spot_f(spot.begin(), spot.end()), drift_f(drift.begin(), drift.end()),
        vols_f(vols.begin(), vols.end()), acc_f(acc.begin(), acc.end()), sched_f(51), ee_f(51);
    for (int i = 0; i < 51; ++i) sched_f[i] = 0.5f * i;
    InterestRateSwap<float> payOff_f(sched_f.data(), sched_f.data(), sched_f.data(), 10.0f,
                                     0.067f, (float)expiry, 0.5f);
    calculateExposureMultiGPU(ee_f.data(), payOff_f, acc_f.data(), spot_f.data(), drift_f.data(),
                              vols_f.data(), 0.0f, num_gpus, scenarios, 0.01f);
    if (hjm_dropin::last_status() != HJM_OK) return 2;
    for (int t = 0; t < 51; ++t) std::printf("%.17g %.9g\n", ee[t], (double)ee_f[t]);
    return 0;
}
