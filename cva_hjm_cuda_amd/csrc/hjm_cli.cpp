// hjm_cli.cpp -- command-line driver over libhjm_cva.so, the successor of the reference's
// test harness (simulation_gpu_testing.cpp:175-259): the same flags and the same printed
// expected-exposure profile ("%1.6f " x 51), so the two can be run side by side.
//
// Reference flags: -single_gpu | -multi_gpu (4 shards) | -cpu, -scenarios N (100000),
// -dt X (0.01), -scale X (accepted, unused). The swap is the reference driver's payer IRS
// (notional 10, K = 6.7%, expiry 25y, dtau 0.5) unless overridden.
// Extensions: -gpus N, -float (the float overload), -market FILE, -notional, -K, -expiry,
// -dtau, -cva (PFE 97.5% and CVA with -hazard, -recovery, -seed, -fp32).
// Without -market the benchmark market is used: Nelson-Siegel curve (3%, -2%, 1%, 2y),
// loadings 0.0064 / quadratic / cubic, their Simpson HJM drift, accrual 0.5.
// -market FILE: whitespace-separated numbers after the labels "spot_rates:" (51),
// "volatilities:" (153, factor-major), "drifts:" (51) and "accrual:" (51).
// -cpu: this library has no CPU path; the driver says so and exits with status 2.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "simulation_gpu.h"

namespace {

bool has_flag(int argc, char** argv, const char* name)
{
    for (int i = 1; i < argc; ++i)
        if (!std::strcmp(argv[i], name)) return true;
    return false;
}

template <typename T>
T flag_value(int argc, char** argv, const char* name, T dflt, bool* bad)
{
    for (int i = 1; i < argc; ++i) {
        if (std::strcmp(argv[i], name)) continue;
        if (i + 1 >= argc) {
            *bad = true;
            return dflt;
        }
        std::istringstream in(argv[i + 1]);
        T v{};
        if (!(in >> v)) {
            *bad = true;
            return dflt;
        }
        return v;
    }
    return dflt;
}

const char* flag_string(int argc, char** argv, const char* name)
{
    for (int i = 1; i + 1 < argc; ++i)
        if (!std::strcmp(argv[i], name)) return argv[i + 1];
    return nullptr;
}

struct MarketData {
    std::vector<double> spot, vols, drifts, accrual;
};

void benchmark_market(MarketData& m)
{
    m.spot.resize(HJM_TENORS);
    m.vols.resize(3 * HJM_TENORS);
    m.drifts.resize(HJM_TENORS);
    m.accrual.assign(HJM_TENORS, 0.5);
    for (int j = 0; j < HJM_TENORS; ++j) {
        const double tau = 0.5 * j, x = tau / 2.0;
        m.spot[j] = 0.03 - 0.02 * std::exp(-x) + 0.01 * x * std::exp(-x);
        m.vols[j] = 0.0064;
        m.vols[HJM_TENORS + j] = -0.0036 + 0.0006 * tau - 2e-5 * tau * tau;
        m.vols[2 * HJM_TENORS + j] = -0.0047 + 0.0018 * tau - 1.5e-4 * tau * tau + 3e-6 * tau * tau * tau;
    }
}

bool read_market(const char* path, MarketData& m)
{
    std::ifstream in(path);
    if (!in) {
        std::fprintf(stderr, "ERROR: cannot open market file %s\n", path);
        return false;
    }
    std::vector<double>* cur = nullptr;
    std::string tok;
    while (in >> tok) {
        if (tok == "spot_rates:") cur = &m.spot;
        else if (tok == "volatilities:") cur = &m.vols;
        else if (tok == "drifts:") cur = &m.drifts;
        else if (tok == "accrual:") cur = &m.accrual;
        else {
            char* end = nullptr;
            const double v = std::strtod(tok.c_str(), &end);
            if (!cur || end == tok.c_str() || *end) {
                std::fprintf(stderr, "ERROR: market file %s: unexpected token '%s'\n", path,
                             tok.c_str());
                return false;
            }
            cur->push_back(v);
        }
    }
    if (m.spot.size() != HJM_TENORS || m.vols.size() != 3 * HJM_TENORS ||
        m.drifts.size() != HJM_TENORS || m.accrual.size() != HJM_TENORS) {
        std::fprintf(stderr, "ERROR: market file %s needs 51 spot_rates, 153 volatilities, "
                             "51 drifts and 51 accrual values\n", path);
        return false;
    }
    return true;
}

template <typename real>
void print_profile(const real* ee)
{
    std::printf("Expected Exposure Profile Double\n");
    for (int t = 0; t < HJM_TENORS; ++t) std::printf("%1.6f ", (double)ee[t]);
    std::printf("\n");
}

}  // namespace

int main(int argc, char** argv)
{
    std::printf("valid arguments [-cpu|-single_gpu|-multi_gpu] -scenarios [number] -dt [0.01 - 0.001] "
                "-scale 1.0\n\n");
    bool bad = false;
    const int scenarios = flag_value<int>(argc, argv, "-scenarios", 100000, &bad);
    const double dt = flag_value<double>(argc, argv, "-dt", 0.01, &bad);
    const double scale = flag_value<double>(argc, argv, "-scale", 0.0, &bad);
    const double notional = flag_value<double>(argc, argv, "-notional", 10.0, &bad);
    const double K = flag_value<double>(argc, argv, "-K", 0.067, &bad);
    const double expiry = flag_value<double>(argc, argv, "-expiry", 25.0, &bad);
    const double dtau = flag_value<double>(argc, argv, "-dtau", 0.5, &bad);
    const double hazard = flag_value<double>(argc, argv, "-hazard", 0.02, &bad);
    const double recovery = flag_value<double>(argc, argv, "-recovery", 0.4, &bad);
    const unsigned long long seed = flag_value<unsigned long long>(argc, argv, "-seed", 1234ull, &bad);
    int num_gpus = has_flag(argc, argv, "-multi_gpu") ? 4 : 1;
    num_gpus = flag_value<int>(argc, argv, "-gpus", num_gpus, &bad);
    if (bad) {
        std::fprintf(stderr, "ERROR: a flag is missing its value or the value does not parse\n");
        return 2;
    }
    if (has_flag(argc, argv, "-cpu")) {
        std::fprintf(stderr, "ERROR: -cpu: this library runs the simulation on the GPU only "
                             "(no CPU path is provided)\n");
        return 2;
    }

    MarketData m;
    const char* mfile = flag_string(argc, argv, "-market");
    if (mfile) {
        if (!read_market(mfile, m)) return 2;
    } else {
        benchmark_market(m);
        const int st = hjm_hjm_drift(m.vols.data(), dtau, m.drifts.data());
        if (st != HJM_OK) {
            std::fprintf(stderr, "ERROR: hjm_hjm_drift: %s\n", hjm_status_string(st));
            return 1;
        }
    }

    std::printf("## scenarios %d simulations per scenarios: %g number of gpus: %d dt: %g\n",
                scenarios, expiry / dt, num_gpus, dt);
    std::printf("## scale %g\n", scale);
    std::printf("## Gpu measurements for num_gpus: %d\n", num_gpus);

    if (has_flag(argc, argv, "-cva")) {
        hjm_market mk{m.spot.data(), m.drifts.data(), m.vols.data(), m.accrual.data()};
        hjm_swap sw{notional, K, expiry, dtau};
        hjm_config cfg{};
        cfg.paths = scenarios;
        cfg.dt = dt;
        cfg.precision = has_flag(argc, argv, "-fp32") ? HJM_FP32_KAHAN : HJM_FP64;
        cfg.rng = HJM_RNG_PHILOX;
        cfg.seed = seed;
        cfg.hazard_rate = hazard;
        cfg.recovery = recovery;
        cfg.pfe_quantile = 0.975;
        double ee[HJM_TENORS], pfe[HJM_TENORS], cva = 0.0, se = 0.0, tm[5];
        hjm_results res{};
        res.expected_exposure = ee;
        res.pfe = pfe;
        res.cva = &cva;
        res.cva_stderr = &se;
        res.timings_ms = tm;
        const int st = hjm_simulate_reference(&mk, &sw, &cfg, &res);
        if (st != HJM_OK) {
            std::fprintf(stderr, "ERROR: hjm_simulate_reference: %s (%d)\n", hjm_status_string(st), st);
            return 1;
        }
        print_profile(ee);
        std::printf("PFE 97.5%% Profile\n");
        for (int t = 0; t < HJM_TENORS; ++t) std::printf("%1.6f ", pfe[t]);
        std::printf("\nCVA %.9g stderr %.3g (hazard %g, recovery %g)\n", cva, se, hazard, recovery);
        std::printf("## device time %.3f ms (paths %.3f ms)\n", tm[4], tm[0]);
        return 0;
    }

    if (has_flag(argc, argv, "-float")) {
        std::vector<float> sp(m.spot.begin(), m.spot.end()), vo(m.vols.begin(), m.vols.end()),
            dr(m.drifts.begin(), m.drifts.end()), ac(m.accrual.begin(), m.accrual.end());
        InterestRateSwap<float> payOff(nullptr, nullptr, nullptr, (float)notional, (float)K,
                                       (float)expiry, (float)dtau);
        float ee[HJM_TENORS];
        calculateExposureMultiGPU(ee, payOff, ac.data(), sp.data(), dr.data(), vo.data(),
                                  (float)scale, num_gpus, scenarios, (float)dt);
        if (hjm_dropin::last_status() != HJM_OK) return 1;
        print_profile(ee);
    } else {
        InterestRateSwap<double> payOff(nullptr, nullptr, nullptr, notional, K, expiry, dtau);
        double ee[HJM_TENORS];
        calculateExposureMultiGPU(ee, payOff, m.accrual.data(), m.spot.data(), m.drifts.data(),
                                  m.vols.data(), scale, num_gpus, scenarios, dt);
        if (hjm_dropin::last_status() != HJM_OK) return 1;
        print_profile(ee);
    }
    return 0;
}
