// hjm_valu.h -- explicit gfx950 VALU instruction forms used by the path kernels.
#pragma once
#include <hip/hip_runtime.h>

namespace hjm {

// ---- the tenor update of the forward path kernels ----------------------------------
// The update of tenor j is exactly six VALU instructions on registers that stay in
// place across steps (t is a temporary; c, 1-c in VGPRs; tenor constants in SGPRs):
//   v_fma  t, c, f[j+1], ddt_j      v_fmac t, v0_j, s0     v_fmac t, v1_j, s1
//   v_fmac t, v2_j, s2              v_fma  f[j], 1-c, f[j], t
//   v_add  acc[j], acc[j], f[j]
// Written out because the compiler otherwise picks the tied v_fmac form for the last
// FMA, which needs a copy per tenor and, with 200+ live values, spills to AGPRs.
__device__ __forceinline__ void op_fma_vvs(double& d, double a, double b, double c)
{
    asm("v_fma_f64 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "s"(c));
}
__device__ __forceinline__ void op_fma_vvs(float& d, float a, float b, float c)
{
    asm("v_fma_f32 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "s"(c));
}
__device__ __forceinline__ void op_fmac_sv(double& d, double a, double b)
{
    asm("v_fmac_f64 %0, %1, %2" : "+v"(d) : "s"(a), "v"(b));
}
__device__ __forceinline__ void op_fmac_sv(float& d, float a, float b)
{
    asm("v_fmac_f32 %0, %1, %2" : "+v"(d) : "s"(a), "v"(b));
}
__device__ __forceinline__ void op_fma_inplace(double& f, double a, double t)
{
    asm("v_fma_f64 %0, %1, %0, %2" : "+v"(f) : "v"(a), "v"(t));
}
__device__ __forceinline__ void op_fma_inplace(float& f, float a, float t)
{
    asm("v_fma_f32 %0, %1, %0, %2" : "+v"(f) : "v"(a), "v"(t));
}
__device__ __forceinline__ void op_add_inplace(double& a, double x)
{
    asm("v_add_f64 %0, %0, %1" : "+v"(a) : "v"(x));
}
__device__ __forceinline__ void op_add_inplace(float& a, float x)
{
    asm("v_add_f32 %0, %0, %1" : "+v"(a) : "v"(x));
}

// ---- adjoint sweep forms ------------------------------------------------------------
// d = a * b + c with all operands in VGPRs
__device__ __forceinline__ void op_fma_vvv(float& d, float a, float b, float c)
{
    asm("v_fma_f32 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "v"(c));
}
// d = a * b with a, b in VGPRs
__device__ __forceinline__ void op_mul_vv(float& d, float a, float b)
{
    asm("v_mul_f32 %0, %1, %2" : "=v"(d) : "v"(a), "v"(b));
}
// d += a * b with a, b in VGPRs, in place
__device__ __forceinline__ void op_fmac_vv(float& d, float a, float b)
{
    asm("v_fmac_f32 %0, %1, %2" : "+v"(d) : "v"(a), "v"(b));
}

}  // namespace hjm
