/*
 * product.h -- the interest-rate-swap description taken by the drop-in entry points.
 *
 * Same type name, template parameter, constructor argument order and member names as
 * the reference's InterestRateSwap<real> (product.h:4-17), so that callers compile
 * unchanged. Of its members the exposure path reads notional, K, expiry and dtau (as
 * the reference does, simulation_gpu.cu:959,1103,1133); the three schedule pointers are
 * carried but unused.
 */
#ifndef HJM_CVA_PRODUCT_H
#define HJM_CVA_PRODUCT_H

template <typename real>
struct InterestRateSwap {
    InterestRateSwap(real* pricing_points_, real* floating_schedule_, real* fixed_schedule_,
                     real notional_, real K_, real expiry_, real dtau_)
        : pricing_points(pricing_points_), realing_schedule(floating_schedule_),
          fixed_schedule(fixed_schedule_), notional(notional_), K(K_), dtau(dtau_),
          expiry(expiry_)
    {
    }

    real* pricing_points;
    real* realing_schedule;  /* floating-leg schedule (the reference's member name) */
    real* fixed_schedule;
    real notional;
    real K;
    real dtau;
    real expiry;
};

#endif
