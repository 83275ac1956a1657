// Device helpers shared by the kernel translation units of libwlsq_hip.
#pragma once
#include "wlsq_internal.h"

namespace wlsq {
namespace {

// acc + a*b (complex); the product's parts are formed with fused multiply-adds.
__device__ __forceinline__ float2 cmac(float2 acc, float2 a, float2 b)
{
    acc.x = fmaf(a.x, b.x, acc.x);
    acc.x = fmaf(-a.y, b.y, acc.x);
    acc.y = fmaf(a.x, b.y, acc.y);
    acc.y = fmaf(a.y, b.x, acc.y);
    return acc;
}

// Re(a * conj(b))
__device__ __forceinline__ float re_mul_conj(float2 a, float2 b)
{
    return fmaf(a.y, b.y, a.x * b.x);
}

// acc + a*b with the rounding of the CPU path: the product's parts are
// rounded separately, (a.x*b.x - a.y*b.y, a.x*b.y + a.y*b.x), then added to
// acc (thrust::complex<float> products in an unconjugated dot,
// cpp_src/mkl_linalg.cpp:7-11). No contraction into fused multiply-adds.
__device__ __forceinline__ float2 cmac_exact(float2 acc, float2 a, float2 b)
{
#pragma clang fp contract(off)
    const float re = a.x * b.x - a.y * b.y;
    const float im = a.x * b.y + a.y * b.x;
    acc.x = acc.x + re;
    acc.y = acc.y + im;
    return acc;
}

// Re(a * conj(b)) = a.x*b.x + a.y*b.y, each product rounded separately
// (cpp_src/extrapolation_cpu.cpp:74-79)
__device__ __forceinline__ float re_mul_conj_exact(float2 a, float2 b)
{
#pragma clang fp contract(off)
    return a.x * b.x + a.y * b.y;
}

// Nearest kappa sample to omega*s on kappa_q = kappa0 + q/inv_dk, ties to the
// lower q (numpy argmin in py_src/operatorsPreparation/util.py:5-8). One
// rounding per operation and no contraction, so the host-side expansion and
// the CPU restatement compute the same q bit for bit. Returns -q:
// q = ceil(u - 0.5) is -floor(-u + 0.5), i.e. minus the round-half-up
// conversion of -u, one instruction (v_cvt_rpi_i32_f32 with a negated
// operand) instead of an add, a ceil and a conversion. For u in [0, nk - 1]
// with nk <= 2^22 the sum -u + 0.5 is exact, so both forms give the same q
// for every float input (checked over all 2^32 of them on gfx950).
__device__ __forceinline__ int kappa_index_neg(float omega, float s, float kappa0, float inv_dk,
                                               int nk)
{
#pragma clang fp contract(off)
    float t = omega * s;
    t = t - kappa0;
    float u = t * inv_dk;
    // clamp to [0, nk-1] in one instruction; a NaN u gives q = 0 (the
    // conversion of NaN is 0), as the rule's "not above 0 -> 0" does
    u = __builtin_amdgcn_fmed3f(u, 0.0f, (float)(nk - 1));
    int r;
    asm("v_cvt_rpi_i32_f32_e64 %0, -%1" : "=v"(r) : "v"(u));
    return r;
}

__device__ __forceinline__ int kappa_index(float omega, float s, float kappa0, float inv_dk, int nk)
{
    return -kappa_index_neg(omega, s, kappa0, inv_dk, nk);
}

// *(base + off bytes): a uniform base and a 32-bit unsigned lane offset, the
// form of a global load with a scalar base (no 64-bit address per lane)
template <typename T>
__device__ __forceinline__ const T &at_bytes(const T *base, unsigned off)
{
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + off);
}

// Position p of a padded row lives at sub-row (p mod R), slot p / R: lanes
// that read or write p = R*lane + c for a fixed c touch consecutive slots.
template <int R>
__device__ __forceinline__ int swz(int p, int Q)
{
    return (p & (R - 1)) * Q + p / R;
}

template <typename K>
hipError_t ensure_lds(K kernel, size_t bytes)
{
    if (bytes <= 64 * 1024)
        return hipSuccess;
    return hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

}  // namespace
}  // namespace wlsq
