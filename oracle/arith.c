/*
 * arith.c -- CPU restatement of the reference Magnitude and Phase kernels (TEST INFRASTRUCTURE:
 * the checker, never the product).
 */
#include "oracle.h"
#include <math.h>
#include <float.h>

/* Magnitude, SIMD body: ago_haf_cpu_arithmetic.cpp:5031-5078.
 * gx^2 and gy^2 as int32, converted to double, added, sqrt_pd, cvtpd_epi32 (round half to even
 * under the default MXCSR), packs_epi32 (saturate to int16). */
static inline int16_t mag_body(int gx, int gy)
{
    double s = (double)(gx * gx) + (double)(gy * gy);
    double r = nearbyint(sqrt(s));
    if (r > 32767.0) r = 32767.0;
    return (int16_t)r;
}

/* Magnitude, postfix columns: ago_haf_cpu_arithmetic.cpp:5087-5091,
 * (vx_int16)round(sqrtf((float)(gx*gx) + (float)(gy*gy))). */
static inline int16_t mag_tail(int gx, int gy)
{
    float t = (float)(gx * gx) + (float)(gy * gy);
    t = sqrtf(t);
    return (int16_t)(int)round(t);
}

int16_t oracle_magnitude_px(int gx, int gy) { return mag_body(gx, gy); }

/* Magnitude_S16_S16S16: ago_haf_cpu_arithmetic.cpp:4989-5097, full frame (rows 0..h-1).
 * Row starts of reference S16 images are 16-byte aligned (32-byte aligned base, ALIGN16 stride,
 * ago_util.cpp:187-200,1107), so the prefix is empty; columns [0, w & ~7) take the SIMD rule and
 * the last w mod 8 columns the scalar postfix rule.
 * Pinned by: bit-exact against oracle/_ref; exhaustive closed form
 * round-half-even(sqrt(n)) == isqrt rule over all (gx, gy) in [-1020, 1020]^2 in tests. */
void oracle_magnitude(int w, int h, int16_t *mag, int mag_stride, const int16_t *gx, int gx_stride,
                      const int16_t *gy, int gy_stride)
{
    int body = w & ~7;
    for (int y = 0; y < h; y++) {
        const int16_t *px = (const int16_t *)((const uint8_t *)gx + (long)y * gx_stride);
        const int16_t *py = (const int16_t *)((const uint8_t *)gy + (long)y * gy_stride);
        int16_t *pm = (int16_t *)((uint8_t *)mag + (long)y * mag_stride);
        for (int x = 0; x < w; x++)
            pm[x] = x < body ? mag_body(px[x], py[x]) : mag_tail(px[x], py[x]);
    }
}

/* atan2 polynomial constants: ago_haf_cpu.h:32-36.  Both factors are float literals
 * (0.9997878412794807f * 57.29577951308232f), so each constant is a rounded float product and the
 * polynomial below is evaluated in float. */
#define P1 (0.9997878412794807f * 57.29577951308232f)
#define P3 (-0.3258083974640975f * 57.29577951308232f)
#define P5 (0.1555786518463281f * 57.29577951308232f)
#define P7 (-0.04432655554792128f * 57.29577951308232f)

/* HafCpu_FastAtan2_deg: ago_haf_cpu_arithmetic.cpp:9564-9589. */
static float fast_atan2_deg(int gx, int gy)
{
    uint16_t ax = (uint16_t)(gx < 0 ? -gx : gx), ay = (uint16_t)(gy < 0 ? -gy : gy);
    float a, c, c2;
    if (ax >= ay) {
        c = (float)ay / ((float)ax + (float)DBL_EPSILON);
        c2 = c * c;
        a = (((P7 * c2 + P5) * c2 + P3) * c2 + P1) * c;
    } else {
        c = (float)ax / ((float)ay + (float)DBL_EPSILON);
        c2 = c * c;
        a = 90.f - (((P7 * c2 + P5) * c2 + P3) * c2 + P1) * c;
    }
    if (gx < 0) a = 180.f - a;
    if (gy < 0) a = 360.f - a;
    return a;
}

/* Phase_U8_S16S16: ago_haf_cpu_arithmetic.cpp:9592-9631 (the #else branch),
 * out = (u8)((u32)(a * (128.f/180.f) + 0.5) & 0xFF), the product in float, the +0.5 in double.
 * Pinned by: bit-exact against oracle/_ref on random gradients and on the exhaustive
 * (gx, gy) in [-1020, 1020]^2 grid (the grid is fed to the reference through its own Phase
 * node in tests); closed form checks at the axes (0, 64, 128, 192 units). */
uint8_t oracle_phase_px(int gx, int gy)
{
    float scale = (float)128 / 180.f;
    float arct = fast_atan2_deg(gx, gy);
    return (uint8_t)((uint32_t)((double)(arct * scale) + 0.5) & 0xFF);
}

void oracle_phase(int w, int h, uint8_t *phase, int phase_stride, const int16_t *gx, int gx_stride,
                  const int16_t *gy, int gy_stride)
{
    for (int y = 0; y < h; y++) {
        const int16_t *px = (const int16_t *)((const uint8_t *)gx + (long)y * gx_stride);
        const int16_t *py = (const int16_t *)((const uint8_t *)gy + (long)y * gy_stride);
        uint8_t *pp = phase + (long)y * phase_stride;
        for (int x = 0; x < w; x++) pp[x] = oracle_phase_px(px[x], py[x]);
    }
}
