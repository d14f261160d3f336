// hjm_rng.h -- counter-based normal generator for the HJM path kernels.
//
// The steps of a run are grouped in date blocks of `stride` steps (dtau / dt). Step n lies
// in block b = n / stride at offset s = n % stride; it belongs to class g = s % 4 and is
// the q-th (q = s / 4) step of that class in the block. The class's shocks form a stream:
// shock j = 3 q + k is factor k of its q-th step. Call c of the stream (path p, block b,
// class g) is one Philox4x32-10 block (Salmon et al., SC'11; constants as published with
// Random123) keyed by the 64-bit seed, with counter (b, p, g, c); Box-Muller turns its
// four 32-bit outputs into the stream's shocks 4c .. 4c + 3:
//   (4c, 4c + 1) = r cos, r sin of (x0, x1);  (4c + 2, 4c + 3) = r cos, r sin of (x2, x3).
// A 50-step block thus costs 38 Philox blocks and 76 Box-Muller pairs (0.76 and 1.52 per
// step) against one block and 1.75 pairs per step when every step draws its own block,
// and a matrix-core lane, which generates the steps of one class, makes every call once.
// The Philox words are exact (the CPU oracle reproduces them bit for bit); Box-Muller runs
// on the hardware's log2 / sqrt / sin / cos instructions, within the bound stated at
// box_muller of the exact transform of the same uniforms.
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

// (z0, z1) = r cos, r sin of the Philox pair (a, b) on the hardware's transcendental
// instructions (gfx950: v_log_f32 = log2, v_sqrt_f32, v_cos_f32 / v_sin_f32 of an angle in
// turns -- exactly the form Box-Muller needs):
//   u = fma(float(a), 2^-32, 2^-33) in (0, 1]    (float(a): a rounded to 24 bits)
//   t = float(b) 2^-32 in [0, 1]                  (turns)
//   r = sqrt(-2 ln 2 * log2(u)),  z0 = r cos(2 pi t),  z1 = r sin(2 pi t)
// u and t are exact functions of the Philox outputs (the CPU oracle computes the same
// floats); r, cos and sin carry the instructions' errors. Measured on an MI355X over every
// distinct u and t: |r - r_exact| <= 1.4e-7 r, |cos - cos_exact|, |sin - sin_exact| <=
// 1.3e-7, so |z - z_exact| <= 2.6e-7 |z| + 8.8e-7 (r <= sqrt(66 ln 2) = 6.77). Four
// transcendental and seven plain VALU instructions per pair.
// The radius of word a and the cosine and sine of word b (the factors of box_muller).
__device__ __forceinline__ float bm_radius(uint32_t a)
{
    const float u = __builtin_fmaf((float)a, 0x1p-32f, 0x1p-33f);
    return __builtin_amdgcn_sqrtf(-0x1.62e430p+0f * __builtin_amdgcn_logf(u));
}
__device__ __forceinline__ void bm_angle(uint32_t b, float& c, float& s)
{
    const float t = (float)b * 0x1p-32f;
    c = __builtin_amdgcn_cosf(t);
    s = __builtin_amdgcn_sinf(t);
}

__device__ __forceinline__ void box_muller(uint32_t a, uint32_t b, float& z0, float& z1)
{
    const float r = bm_radius(a);
    float c, s;
    bm_angle(b, c, s);
    z0 = r * c;
    z1 = r * s;
}

// Shocks 4 call .. 4 call + 3 of the stream (path, block, cls).
__device__ __forceinline__ void stream_call(uint32_t seed_lo, uint32_t seed_hi, uint32_t path,
                                            uint32_t block, uint32_t cls, uint32_t call,
                                            float (&z)[4])
{
    uint32_t x[4];
    philox4x32_10(block, path, cls, call, seed_lo, seed_hi, x);
    box_muller(x[0], x[1], z[0], z[1]);
    box_muller(x[2], x[3], z[2], z[3]);
}

// The three factor shocks of step n, computed directly: one call when they lie in one
// call's four shocks (q % 4 = 0 or 3), two otherwise.
__device__ __forceinline__ Normals3 step_normals(uint32_t seed_lo, uint32_t seed_hi,
                                                 uint32_t path, uint32_t n, uint32_t stride)
{
    const uint32_t b = n / stride, s = n % stride, g = s & 3u, j0 = 3u * (s >> 2);
    const uint32_t c0 = j0 >> 2;
    float a[4], c[4];
    // one branch per position of the step's first shock in its call (uniform across a
    // wave): each keeps only the Box-Muller outputs it returns
    switch (j0 & 3u) {
    case 0u:
        stream_call(seed_lo, seed_hi, path, b, g, c0, a);
        return Normals3{a[0], a[1], a[2]};
    case 1u:
        stream_call(seed_lo, seed_hi, path, b, g, c0, a);
        return Normals3{a[1], a[2], a[3]};
    case 2u:
        stream_call(seed_lo, seed_hi, path, b, g, c0, a);
        stream_call(seed_lo, seed_hi, path, b, g, c0 + 1u, c);
        return Normals3{a[2], a[3], c[0]};
    default:
        stream_call(seed_lo, seed_hi, path, b, g, c0, a);
        stream_call(seed_lo, seed_hi, path, b, g, c0 + 1u, c);
        return Normals3{a[3], c[0], c[1]};
    }
}

// Steps [16 m, 16 m + 16) of a date block are quads 4 m .. 4 m + 3 of every class, i.e.
// calls 3 m .. 3 m + 2 of the four class streams: 12 calls, 48 shocks, all used. Kernels
// that walk a block step by step (one path per lane) stage a group's shocks in LDS, in a
// column of their own (rows `ld` floats apart), and read each step's three back: 0.75
// calls per step and nothing carried in registers across steps. Steps past the last
// whole group of a block (2 of 50) use step_normals.
constexpr int kGroupSteps = 16;
constexpr int kGroupShocks = 48;

__device__ __forceinline__ void fill_group(uint32_t seed_lo, uint32_t seed_hi, uint32_t path,
                                           uint32_t block, int m, float* col, int ld)
{
    // one call at a time (the caller's loop holds the path state in registers)
#pragma unroll 1
    for (int g = 0; g < 4; ++g)
#pragma unroll 1
        for (int cc = 0; cc < 3; ++cc) {
            float v[4];
            stream_call(seed_lo, seed_hi, path, block, (uint32_t)g, (uint32_t)(3 * m + cc), v);
#pragma unroll
            for (int e = 0; e < 4; ++e) col[(12 * g + 4 * cc + e) * ld] = v[e];
        }
}

// shocks of step s of the staged group: class s % 4, quad (s / 4) % 4 of the group
__device__ __forceinline__ Normals3 group_step(const float* col, int ld, int s)
{
    const int r = 12 * (s & 3) + 3 * ((s >> 2) & 3);
    return Normals3{col[r * ld], col[(r + 1) * ld], col[(r + 2) * ld]};
}

}  // namespace hjm
