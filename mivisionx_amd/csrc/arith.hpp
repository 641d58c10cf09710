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
// path bit for bit).
MVX_DEV int32_t mag_body(int32_t gx, int32_t gy)
{
    uint32_t s = (uint32_t)(gx * gx) + (uint32_t)(gy * gy);   // <= 2^31
    uint32_t k = (uint32_t)sqrtf((float)s);
    while ((uint64_t)k * k > s) k--;
    while ((uint64_t)(k + 1) * (k + 1) <= s) k++;
    uint32_t r = (s - k * k > k) ? k + 1 : k;
    return (int32_t)min(r, 32767u);
}

// Magnitude postfix rule (columns w & ~7 .. w-1), ago_haf_cpu_arithmetic.cpp:5087-5091:
// (vx_int16)round(sqrtf((float)(gx*gx) + (float)(gy*gy))), float adds and sqrt correctly rounded.
MVX_DEV int32_t mag_tail(int32_t gx, int32_t gy)
{
    float t = (float)(gx * gx) + (float)(gy * gy);
    t = __fsqrt_rn(t);
    return (int32_t)(int16_t)(int32_t)roundf(t);
}

// atan2 polynomial of ago_haf_cpu.h:32-36: each constant is a product of two float literals,
// rounded to float, and the polynomial runs in float.
constexpr float ATAN_P1 = 0.9997878412794807f * 57.29577951308232f;
constexpr float ATAN_P3 = -0.3258083974640975f * 57.29577951308232f;
constexpr float ATAN_P5 = 0.1555786518463281f * 57.29577951308232f;
constexpr float ATAN_P7 = -0.04432655554792128f * 57.29577951308232f;

// HafCpu_FastAtan2_deg (ago_haf_cpu_arithmetic.cpp:9564-9589) and the Phase quantisation (:9623-9626)
MVX_DEV uint32_t phase_px(int32_t gx, int32_t gy)
{
    const uint32_t ax = (uint16_t)(gx < 0 ? -gx : gx), ay = (uint16_t)(gy < 0 ? -gy : gy);
    const float eps = (float)2.2204460492503131e-16;   // (float)DBL_EPSILON
    float a, c, c2;
    if (ax >= ay) {
        c = __fdiv_rn((float)ay, (float)ax + eps);
        c2 = c * c;
        a = (((ATAN_P7 * c2 + ATAN_P5) * c2 + ATAN_P3) * c2 + ATAN_P1) * c;
    } else {
        c = __fdiv_rn((float)ax, (float)ay + eps);
        c2 = c * c;
        a = 90.f - (((ATAN_P7 * c2 + ATAN_P5) * c2 + ATAN_P3) * c2 + ATAN_P1) * c;
    }
    if (gx < 0) a = 180.f - a;
    if (gy < 0) a = 360.f - a;
    const float scale = 128.f / 180.f;
    return (uint32_t)((double)(a * scale) + 0.5) & 0xffu;
}

}  // namespace mvx
