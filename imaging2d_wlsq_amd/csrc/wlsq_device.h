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

// Exact image accumulation. A finite float32 v is m * 2^(e - 150) with m < 2^24
// (e the biased exponent, subnormals as e = 1 without the hidden bit), i.e.
// m * 2^p units of 2^-149 with p = max(e, 1) - 1 in [0, 253]. Word k of a
// point's accumulator holds multiples of 2^(16k) units, so v adds m << (p mod 16)
// (below 2^39) to word p / 16 as a 64-bit integer: 16 words cover every finite
// float, integer addition is exact and independent of order, and 2^24 terms fit
// in a word without overflow. Word 16 holds flags: bit 0 NaN, bit 1 +inf,
// bit 2 -inf. The words of a point lie `stride` apart.
__device__ __forceinline__ void exact_add(unsigned long long *acc, long long stride, long long i,
                                          float v)
{
    const unsigned u = __float_as_uint(v);
    const unsigned e = (u >> 23) & 0xffu;
    unsigned m = u & 0x7fffffu;
    if (e == 0xffu) {
        const unsigned long long f = m ? 1ull : ((u >> 31) ? 4ull : 2ull);
        atomicOr(acc + kExactWords * stride + i, f);
        return;
    }
    if (e == 0u && m == 0u)
        return;  // +-0 adds nothing
    int p = 0;
    if (e) {
        m |= 0x800000u;
        p = (int)e - 1;
    }
    const unsigned long long mag = (unsigned long long)m << (p & 15);
    atomicAdd(acc + (long long)(p >> 4) * stride + i, (u >> 31) ? 0ull - mag : mag);
}

// The float32 nearest (ties to even) to the exact sum held in words w[0..15]
// and flags f of one accumulator (see exact_add); +-inf past the float range,
// NaN for a NaN term or for both infinities. Integer arithmetic only, so the
// host and the device agree bit for bit.
__host__ __device__ inline float exact_round(const unsigned long long *w, unsigned long long f)
{
    if (f & 1ull || (f & 6ull) == 6ull)
        return __builtin_nanf("");
    if (f & 2ull)
        return __builtin_inff();
    if (f & 4ull)
        return -__builtin_inff();
    // carry into 16-bit digits; the signed remainder above word 15 is `top`
    unsigned d[16];
    __int128 carry = 0;
    for (int k = 0; k < 16; ++k) {
        const __int128 t = (__int128)(long long)w[k] + carry;
        d[k] = (unsigned)(t & 0xffff);
        carry = t >> 16;  // arithmetic shift: floor division by 2^16
    }
    bool neg = carry < 0;
    unsigned long long top = 0;
    if (neg) {
        // two's complement of (carry, d[15..0]): invert and add one
        unsigned c = 1;
        for (int k = 0; k < 16; ++k) {
            const unsigned t = (~d[k] & 0xffffu) + c;
            d[k] = t & 0xffffu;
            c = t >> 16;
        }
        top = (unsigned long long)(~carry) + c;
    } else {
        top = (unsigned long long)carry;
    }
    // magnitude as 32-bit words W[0..9] (bit 0 of W[0] weighs 2^-149)
    unsigned W[10];
    for (int j = 0; j < 8; ++j)
        W[j] = d[2 * j] | (d[2 * j + 1] << 16);
    W[8] = (unsigned)(top & 0xffffffffull);
    W[9] = (unsigned)(top >> 32);
    int jt = 9;
    while (jt >= 0 && W[jt] == 0)
        --jt;
    if (jt < 0)
        return 0.0f;
    // the leading 64 bits and whether any bit below them is set
    unsigned long long hi = W[jt];
    int e0 = 0;  // weight of hi's bit 0, in bits above 2^-149
    bool sticky = false;
    if (jt >= 1) {
        hi = (hi << 32) | W[jt - 1];
        e0 = 32 * (jt - 1);
        for (int j = 0; j < jt - 1; ++j)
            sticky = sticky || W[j] != 0;
    }
    const int b = 63 - __builtin_clzll(hi);
    unsigned long long mant = hi;
    int ex = e0;
    if (b > 23) {
        const int sh = b - 23;
        mant = hi >> sh;
        const bool r = (hi >> (sh - 1)) & 1ull;
        const bool st = sticky || (hi & ((1ull << (sh - 1)) - 1ull)) != 0;
        if (r && (st || (mant & 1ull)))
            ++mant;
        ex += sh;
        if (mant >> 24) {
            mant >>= 1;
            ++ex;
        }
    }
    unsigned bits;
    if (mant < (1ull << 23)) {
        bits = (unsigned)mant;  // subnormal (ex is 0 here), exact
    } else {
        const int biased = ex + 1;  // value = mant * 2^(ex - 149), mant in [2^23, 2^24)
        bits = biased >= 255 ? 0x7f800000u : ((unsigned)biased << 23) | ((unsigned)mant & 0x7fffffu);
    }
    if (neg)
        bits |= 0x80000000u;
    float out;
    __builtin_memcpy(&out, &bits, 4);
    return out;
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
