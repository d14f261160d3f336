// hjm_rng.h -- counter-based normal generator for the HJM path kernels.
//
// Step n of Monte Carlo path p draws its three factor shocks from one Philox4x32-10
// block (Salmon et al., SC'11; constants as published with Random123) keyed by the
// 64-bit seed, with counter (n, p, 0, 0). Box-Muller turns the four 32-bit outputs into
// three N(0,1) floats: z0, z1 = r cos, r sin of (x0, x1); z2 = r' cos of (x2, x3). ln,
// sin and cos are fixed FP32 polynomials evaluated with explicit FMAs (coefficients from
// rng_poly_fit.py), so host and GPU produce the same bits; the square root is IEEE-rounded on
// both.
//
// Counter addressing makes any (path, step) directly reachable: paths shard across
// devices without stream offsets, and the adjoint sweep regenerates shocks in reverse.
// This replaces the reference's materialised cuRAND XORWOW normals
// (simulation_gpu.cu:193-216, :1009-1013; simulation_cpu.h:60-90).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hjm {

#define HJM_HD __host__ __device__ __forceinline__

struct Normals3 {
    float z0, z1, z2;
};

// a ^ b ^ c; one v_bitop3_b32 (truth table 0x96) on the GPU
HJM_HD uint32_t xor3(uint32_t a, uint32_t b, uint32_t c)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
#else
    return a ^ b ^ c;
#endif
}

HJM_HD void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                          uint32_t k1, uint32_t out[4])
{
    constexpr uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    constexpr uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)M0 * c0;
        const uint64_t p1 = (uint64_t)M1 * c2;
        const uint32_t n0 = xor3((uint32_t)(p1 >> 32), c1, k0);
        const uint32_t n2 = xor3((uint32_t)(p0 >> 32), c3, k1);
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
        k0 += W0;
        k1 += W1;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

HJM_HD float bits_to_float(uint32_t b) { return __builtin_bit_cast(float, b); }
HJM_HD uint32_t float_to_bits(float f) { return __builtin_bit_cast(uint32_t, f); }

// ln(u), u in [2^-24, 1]: u = 2^e m with m in [sqrt(1/2), sqrt(2)], ln m = x + x^2 P(x).
HJM_HD float rng_log(float u)
{
#pragma clang fp contract(off)
    const uint32_t b = float_to_bits(u);
    int e = (int)((b >> 23) & 0xffu) - 127;
    uint32_t mb = (b & 0x007fffffu) | 0x3f800000u;
    if (mb > 0x3fb504f3u) {
        mb -= 0x00800000u;
        e += 1;
    }
    const float x = bits_to_float(mb) - 1.0f;
    float p = -0x1.383d96p-4f;
    p = __builtin_fmaf(p, x, 0x1.084ep-3f);
    p = __builtin_fmaf(p, x, -0x1.0f375cp-3f);
    p = __builtin_fmaf(p, x, 0x1.22709cp-3f);
    p = __builtin_fmaf(p, x, -0x1.54274ap-3f);
    p = __builtin_fmaf(p, x, 0x1.99a402p-3f);
    p = __builtin_fmaf(p, x, -0x1.000428p-2f);
    p = __builtin_fmaf(p, x, 0x1.55555p-2f);
    p = __builtin_fmaf(p, x, -0x1.fffff8p-2f);
    const float x2 = x * x;
    const float lnm = __builtin_fmaf(x2, p, x);
    const float fe = (float)e;
    return __builtin_fmaf(fe, 0x1.62e4p-1f, __builtin_fmaf(fe, 0x1.7f7d1ep-20f, lnm));
}

// sin and cos of 2 pi k / 2^24 for k in [0, 2^24), quadrant reduction in integers.
HJM_HD void rng_sincos_turn24(uint32_t k24, float& s, float& c)
{
#pragma clang fp contract(off)
    const uint32_t q = (k24 + (1u << 21)) >> 22;
    const int32_t ri = (int32_t)k24 - (int32_t)(q << 22);
    const float r = (float)ri * 0x1p-24f;
    const float r2 = r * r;
    float ps = 0x1.4bc89ap+5f;
    ps = __builtin_fmaf(ps, r2, -0x1.32ca9ep+6f);
    ps = __builtin_fmaf(ps, r2, 0x1.466bbap+6f);
    ps = __builtin_fmaf(ps, r2, -0x1.4abbcep+5f);
    ps = __builtin_fmaf(ps, r2, 0x1.921fb6p+2f);
    const float sn = r * ps;
    float pc = 0x1.d9d57cp+5f;
    pc = __builtin_fmaf(pc, r2, -0x1.55c5ep+6f);
    pc = __builtin_fmaf(pc, r2, 0x1.03c1dep+6f);
    pc = __builtin_fmaf(pc, r2, -0x1.3bd3ccp+4f);
    const float cs = __builtin_fmaf(r2, pc, 1.0f);
    const uint32_t qq = q & 3u;
    const float a = (qq & 1u) ? cs : sn;   // sin of the full angle, up to sign
    const float bb = (qq & 1u) ? sn : cs;  // cos of the full angle, up to sign
    s = (qq & 2u) ? -a : a;
    c = ((qq + 1u) & 2u) ? -bb : bb;
}

HJM_HD float rng_cos_turn24(uint32_t k24)
{
    float s, c;
    rng_sincos_turn24(k24, s, c);
    return c;
}

// IEEE-rounded square root for x == 0 and x >= 2^-96 (the Box-Muller radii lie in
// [0, 33.3]): on the GPU, v_sqrt_f32 (within 1 ulp) and a one-ulp correction from the two
// neighbours' residuals, the hardware-rate form of the correctly rounded expansion
// without its tiny-argument rescaling.
HJM_HD float rng_sqrt(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
#pragma clang fp contract(off)
    const float s = __builtin_amdgcn_sqrtf(x);
    const float dn = bits_to_float(float_to_bits(s) - 1u);
    const float up = bits_to_float(float_to_bits(s) + 1u);
    const float rdn = __builtin_fmaf(-dn, s, x);
    const float rup = __builtin_fmaf(-up, s, x);
    const float t = (rdn <= 0.0f) ? dn : s;
    return (rup > 0.0f) ? up : t;
#else
    return __builtin_sqrtf(x);
#endif
}

HJM_HD Normals3 step_normals(uint32_t seed_lo, uint32_t seed_hi, uint32_t path, uint32_t step)
{
#pragma clang fp contract(off)
    uint32_t x[4];
    philox4x32_10(step, path, 0u, 0u, seed_lo, seed_hi, x);
    const float u0 = (float)((x[0] >> 8) + 1u) * 0x1p-24f;
    const float u2 = (float)((x[2] >> 8) + 1u) * 0x1p-24f;
    const float r0 = rng_sqrt(-2.0f * rng_log(u0));
    const float r2 = rng_sqrt(-2.0f * rng_log(u2));
    float s0, c0;
    rng_sincos_turn24(x[1] >> 8, s0, c0);
    const float c2 = rng_cos_turn24(x[3] >> 8);
    Normals3 z;
    z.z0 = r0 * c0;
    z.z1 = r0 * s0;
    z.z2 = r2 * c2;
    return z;
}

}  // namespace hjm
