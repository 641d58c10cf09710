// arith.hpp -- per-pixel Magnitude and Phase rules of the reference CPU path, shared by the
// per-node kernels (arith.hip) and the fused batched chain (chain_batched.hip).
// Built with -ffp-contract=off and correctly rounded fp32 divide/sqrt so that float sequences
// follow the reference CPU code operation for operation.
#pragma once
#include "mvx_common.hpp"

namespace mvx {

// Magnitude SIMD-body rule of ago_haf_cpu_arithmetic.cpp:5031-5078: cvtpd_epi32(sqrt_pd(gx^2+gy^2))
// (round half to even), packs_epi32 saturation.  Evaluated exactly in integers: with
// k = floor(sqrt(s)), round(sqrt(s)) is k+1 iff s - k^2 > k (sqrt(s) is never a half-integer for
// integer s, and the gap to k+1/2 is far above double rounding error, so this equals the double
// path bit for bit).  s <= 2^31, so sqrt(s) <= 46341: the hardware square root of (float)s (error
// well under 1) truncates to within 1 of k, one correction each way recovers k, and every square
// involved stays below 2^32 with factors below 2^24 (v_mul_u32_u24).
MVX_DEV int32_t mag_body(int32_t gx, int32_t gy)
{
    const uint32_t s = (uint32_t)__mul24(gx, gx) + (uint32_t)__mul24(gy, gy);   // |g| <= 32768
    uint32_t k = (uint32_t)__builtin_amdgcn_sqrtf((float)s);
    const uint32_t kk = __umul24(k, k);
    if (kk > s) k--;
    else if (kk + 2 * k + 1 <= s) k++;
    const uint32_t r = (s - __umul24(k, k) > k) ? k + 1 : k;
    return (int32_t)min(r, 32767u);
}

// Canny L2 magnitude of the reference's SSE loops (ago_haf_cpu_canny.cpp:646-1056): _mm_sqrt_ps of the
// float sum, _mm_cvtps_epi32 (round half to even) and the unsigned 16-bit pack.  f >= 0 is an
// integer-valued float below 2^32; sqrtf is the correctly rounded square root under this build's
// -fhip-fp32-correctly-rounded-divide-sqrt, and rint the nearest-even conversion: the same operations
// (checked for every such float by an exhaustive GPU probe).
MVX_DEV int l2_pack_simd(float f)
{
    const int k = __float2int_rn(sqrtf(f));
    return k > 65535 ? 65535 : k;
}

// Magnitude postfix rule (columns w & ~7 .. w-1), ago_haf_cpu_arithmetic.cpp:5087-5091:
// (vx_int16)round(sqrtf((float)(gx*gx) + (float)(gy*gy))), float adds and sqrt correctly rounded.
MVX_DEV int32_t mag_tail(int32_t gx, int32_t gy)
{
    float t = (float)(gx * gx) + (float)(gy * gy);
    t = __fsqrt_rn(t);
    return (int32_t)(int16_t)(int32_t)roundf(t);
}

// Magnitude of the merged Sobel kernels' scalar prefix/postfix columns (ago_haf_cpu_filter.cpp:2496-2499,
// :4004-4007): (vx_int16)sqrtf((float)(gx*gx) + (float)(gy*gy)), truncated rather than rounded.  The sum is
// an integer below 2^24, so the float adds are exact and the correctly rounded square root is the same
// float the reference computes.
MVX_DEV int32_t mag_trunc(int32_t gx, int32_t gy)
{
    float t = (float)(gx * gx) + (float)(gy * gy);
    t = __fsqrt_rn(t);
    return (int32_t)(int16_t)(int32_t)t;
}

// atan2 polynomial of ago_haf_cpu.h:32-36: each constant is a product of two float literals,
// rounded to float, and the polynomial runs in float.
constexpr float ATAN_P1 = 0.9997878412794807f * 57.29577951308232f;
constexpr float ATAN_P3 = -0.3258083974640975f * 57.29577951308232f;
constexpr float ATAN_P5 = 0.1555786518463281f * 57.29577951308232f;
constexpr float ATAN_P7 = -0.04432655554792128f * 57.29577951308232f;

// Correctly rounded a / b for integers 0 <= a <= b <= 32768 (and b = (float)DBL_EPSILON with a = 0):
// hardware reciprocal, product, exact remainder by FMA, one FMA correction.  The sequence is not
// correctly rounded for every float pair; for this domain it agrees with IEEE division on every
// one of the 5.4e8 pairs (checked exhaustively on the GPU by the test suite).
MVX_DEV float div_small_rn(float a, float b)
{
    const float y = __builtin_amdgcn_rcpf(b);
    const float q = a * y;
    const float r = fmaf(-q, b, a);
    return fmaf(r, y, q);
}

// HafCpu_FastAtan2_deg (ago_haf_cpu_arithmetic.cpp:9564-9589) and the Phase quantisation (:9623-9626).
// Both branches divide the smaller magnitude by the larger (+DBL_EPSILON, which only matters when
// both are 0) and evaluate the same polynomial, so one division serves both.  The final
// (u32)((double)v + 0.5) for v = a*scale in [0, 256] is floor(v + 1/2), computed exactly as
// trunc(v) + (v - trunc(v) >= 1/2) (the subtraction is exact).
MVX_DEV uint32_t phase_px(int32_t gx, int32_t gy)
{
    const uint32_t ax = (uint16_t)(gx < 0 ? -gx : gx), ay = (uint16_t)(gy < 0 ? -gy : gy);
    const float eps = (float)2.2204460492503131e-16;   // (float)DBL_EPSILON
    const bool xMajor = ax >= ay;
    const float c = div_small_rn((float)(xMajor ? ay : ax), (float)(xMajor ? ax : ay) + eps);
    const float c2 = c * c;
    const float p = (((ATAN_P7 * c2 + ATAN_P5) * c2 + ATAN_P3) * c2 + ATAN_P1) * c;
    float a = xMajor ? p : 90.f - p;
    if (gx < 0) a = 180.f - a;
    if (gy < 0) a = 360.f - a;
    const float scale = 128.f / 180.f;
    const float v = a * scale;
    const float t = truncf(v);
    return ((uint32_t)t + (v - t >= 0.5f ? 1u : 0u)) & 0xffu;
}

}  // namespace mvx
