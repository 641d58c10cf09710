/*
 * filter.c -- CPU restatement of the reference 3x3 U8 filters (TEST INFRASTRUCTURE: the checker,
 * never the product).  Memory model: see oracle.h.
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ALIGN16_C(x) (((x) + 15) & ~15)

/* Linear-address read with the reference's zero guard outside [0, h*stride). */
static inline int rd8(const uint8_t *p, long n, long i) { return (i < 0 || i >= n) ? 0 : p[i]; }

/* Gaussian3x3: ago_haf_cpu_filter.cpp:1518-1864 (active branch :1681-1861, USE_AVX=0),
 * called for rows 1..h-2 by ago_kernel_api.cpp:15274-15339.
 *   h(r) = p(x-1,r) + 2 p(x,r) + p(x+1,r)          (u16, :1706-1707)
 *   out  = (h(y-1) + 2 h(y) + h(y+1)) >> 4          (:1782-1784)
 * Pinned by: bit-exact against oracle/_ref on random, constant and ramp images, including the
 * wrap columns; constant image => output == input (closed form). */
void oracle_gaussian3x3(int w, int h, uint8_t *dst, int dst_stride, const uint8_t *src, int src_stride)
{
    long n = (long)h * src_stride;
    for (int y = 1; y < h - 1; y++) {
        for (int x = 0; x < w; x++) {
            unsigned acc = 0;
            for (int dy = -1; dy <= 1; dy++) {
                long b = (long)(y + dy) * src_stride + x;
                unsigned hr = rd8(src, n, b - 1) + 2u * rd8(src, n, b) + rd8(src, n, b + 1);
                acc += (dy == 0) ? 2u * hr : hr;
            }
            dst[(long)y * dst_stride + x] = (uint8_t)(acc >> 4);
        }
    }
}

/* Box3x3: ago_haf_cpu_filter.cpp:34-230.  SIMD body mulhi(sum, 7282) (:61,:184) and scalar
 * (char)((float)sum / 9.0f) (:145,:202) both equal floor(sum/9) for sum <= 2295.
 * Pinned by: bit-exact against oracle/_ref; exhaustive check of mulhi == floor(sum/9) over
 * sum in [0, 2295] in tests; constant image => output == input. */
void oracle_box3x3(int w, int h, uint8_t *dst, int dst_stride, const uint8_t *src, int src_stride)
{
    long n = (long)h * src_stride;
    for (int y = 1; y < h - 1; y++) {
        for (int x = 0; x < w; x++) {
            unsigned s = 0;
            for (int dy = -1; dy <= 1; dy++) {
                long b = (long)(y + dy) * src_stride + x;
                s += rd8(src, n, b - 1) + rd8(src, n, b) + rd8(src, n, b + 1);
            }
            dst[(long)y * dst_stride + x] = (uint8_t)(s / 9u);
        }
    }
}

/* Median3x3: ago_haf_cpu_filter.cpp:3842-3945 (qsort in prefix/postfix, a compare-swap network
 * in the body; both give the exact 5th order statistic of the 9 neighbours).
 * Pinned by: bit-exact against oracle/_ref; brute-force sort on tiny inputs. */
void oracle_median3x3(int w, int h, uint8_t *dst, int dst_stride, const uint8_t *src, int src_stride)
{
    long n = (long)h * src_stride;
    for (int y = 1; y < h - 1; y++) {
        for (int x = 0; x < w; x++) {
            int hist[256];
            memset(hist, 0, sizeof(hist));
            for (int dy = -1; dy <= 1; dy++) {
                long b = (long)(y + dy) * src_stride + x;
                hist[rd8(src, n, b - 1)]++;
                hist[rd8(src, n, b)]++;
                hist[rd8(src, n, b + 1)]++;
            }
            int c = 0, v = 0;
            for (v = 0; v < 256; v++) { c += hist[v]; if (c >= 5) break; }
            dst[(long)y * dst_stride + x] = (uint8_t)v;
        }
    }
}

/* Sobel3x3 GX/GY: ago_haf_cpu_filter.cpp:1872 (GX), :2047 (GY), :2237-2460 (GXY, merged form);
 * wrappers ago_kernel_api.cpp:16111-16315 (rows 1..h-2).
 *   gx = (p(x+1,y-1)-p(x-1,y-1)) + 2(p(x+1,y)-p(x-1,y)) + (p(x+1,y+1)-p(x-1,y+1))
 *   gy = [1 2 1].row(y+1) - [1 2 1].row(y-1)                     (int16, :2278-2283, :2434-2436)
 * Pinned by: bit-exact against oracle/_ref; linear ramp p = a*x + b*y => gx = 8a, gy = 8b
 * away from the wrap columns (closed form). */
void oracle_sobel3x3(int w, int h, int16_t *gx, int gx_stride, int16_t *gy, int gy_stride,
                     const uint8_t *src, int src_stride)
{
    long n = (long)h * src_stride;
    for (int y = 1; y < h - 1; y++) {
        for (int x = 0; x < w; x++) {
            long bm = (long)(y - 1) * src_stride + x, b0 = (long)y * src_stride + x, bp = (long)(y + 1) * src_stride + x;
            int dxm = rd8(src, n, bm + 1) - rd8(src, n, bm - 1);
            int dx0 = rd8(src, n, b0 + 1) - rd8(src, n, b0 - 1);
            int dxp = rd8(src, n, bp + 1) - rd8(src, n, bp - 1);
            int hm = rd8(src, n, bm - 1) + 2 * rd8(src, n, bm) + rd8(src, n, bm + 1);
            int hp = rd8(src, n, bp - 1) + 2 * rd8(src, n, bp) + rd8(src, n, bp + 1);
            if (gx) *(int16_t *)((uint8_t *)gx + (long)y * gx_stride + 2L * x) = (int16_t)(dxm + 2 * dx0 + dxp);
            if (gy) *(int16_t *)((uint8_t *)gy + (long)y * gy_stride + 2L * x) = (int16_t)(hp - hm);
        }
    }
}

/* ---- merged Sobel -> Magnitude / Phase ------------------------------------------------------
 * The reference's graph optimizer replaces Sobel_S16S16_U8_3x3_GXY + Magnitude + Phase by one
 * kernel when gx and gy are virtual and no other node reads them (merge rules
 * ago_drama_remove.cpp:155-182, SOLITARY arguments; applied unless the graph's optimizer flags
 * hold NO_NODE_MERGE, :1455).  The wrappers call the merged kernels for rows 1..h-2
 * (ago_kernel_api.cpp:15958-16110), so rows 0 and h-1 of the outputs are never written. */

/* Magnitude of the merged kernels, SSE body: gx^2 + gy^2 in int32, sqrt_pd, cvtpd_epi32 (round
 * half to even), packus_epi32 (saturate to [0, 65535]), stored as int16
 * (ago_haf_cpu_filter.cpp:2602-2653 and :4130-4181). */
static inline int16_t merged_mag_body(int gx, int gy)
{
    double r = nearbyint(sqrt((double)(gx * gx) + (double)(gy * gy)));
    if (r > 65535.0) r = 65535.0;
    return (int16_t)(uint16_t)r;
}

/* Magnitude of the merged kernels, scalar prefix/postfix columns: (vx_int16)sqrtf((float)(gx*gx) +
 * (float)(gy*gy)), i.e. TRUNCATED (ago_haf_cpu_filter.cpp:2496-2499, :2664-2667, :4004-4007,
 * :4192-4195), unlike the Magnitude node's rounded postfix. */
static inline int16_t merged_mag_tail(int gx, int gy)
{
    float m = (float)(gx * gx) + (float)(gy * gy);
    m = sqrtf(m);
    return (int16_t)m;
}

int16_t oracle_merged_magnitude_px(int gx, int gy, int tail) { return tail ? merged_mag_tail(gx, gy) : merged_mag_body(gx, gy); }

/* HafCpu_SobelMagnitudePhase_S16U8_U8_3x3 (ago_haf_cpu_filter.cpp:3972-4202) and
 * HafCpu_SobelMagnitude_S16_U8_3x3 (:2468-2682); mag or phase may be NULL (SobelMagnitude: phase
 * NULL).  Output rows 1..h-2, all columns, 3x3 neighbours at linear addresses as in the Sobel node.
 * The magnitude row start is 16-byte aligned in the reference layout (32-byte aligned buffer,
 * ALIGN16 stride), so the prefix loop is empty, the 16-pixel SSE body covers columns
 * [0, w & ~15) and the scalar postfix the last w mod 16 columns.  The phase is HafCpu_FastAtan2_deg
 * quantised as in the Phase node (:4061-4075, :4196-4197).
 * Pinned by: bit-exact against oracle/_ref graphs whose gx/gy are virtual (widths 1920, 1928, 333,
 * 17, ...); closed form: body columns round half to even, postfix columns truncate, at pixels with
 * a known gradient. */
void oracle_sobel_magnitude_phase(int w, int h, int16_t *mag, int mag_stride, uint8_t *phase, int phase_stride,
                                  const uint8_t *src, int src_stride)
{
    long n = (long)h * src_stride;
    const int body = w & ~15;
    for (int y = 1; y < h - 1; y++) {
        for (int x = 0; x < w; x++) {
            long bm = (long)(y - 1) * src_stride + x, b0 = (long)y * src_stride + x, bp = (long)(y + 1) * src_stride + x;
            int gx = (rd8(src, n, bm + 1) - rd8(src, n, bm - 1)) + 2 * (rd8(src, n, b0 + 1) - rd8(src, n, b0 - 1)) +
                     (rd8(src, n, bp + 1) - rd8(src, n, bp - 1));
            int gy = (rd8(src, n, bp - 1) + 2 * rd8(src, n, bp) + rd8(src, n, bp + 1)) -
                     (rd8(src, n, bm - 1) + 2 * rd8(src, n, bm) + rd8(src, n, bm + 1));
            if (mag) *(int16_t *)((uint8_t *)mag + (long)y * mag_stride + 2L * x) = x < body ? merged_mag_body(gx, gy) : merged_mag_tail(gx, gy);
            if (phase) phase[(long)y * phase_stride + x] = oracle_phase_px(gx, gy);
        }
    }
}

/* HafCpu_SobelPhase_U8_U8_3x3 (ago_haf_cpu_filter.cpp:3946-3970): Sobel GXY into the node's scratch,
 * then Phase_U8_S16S16 from it.  The scratch rows are dstride = ALIGN16(w) BYTES apart (:3958,
 * :3965) although a row holds w int16 values (2w bytes), so when the GXY kernel writes rows
 * top to bottom each row overwrites the second half of the row before it; Phase then reads the
 * rows back with the same stride.  The scratch is emulated literally: call row r (output row r+1)
 * is written at int16 index r * dstride/2, in row order, and read back from there.
 * Pinned by: bit-exact against oracle/_ref graphs whose gx/gy are virtual and read only by a
 * Phase node; closed form: for x >= ALIGN16(w)/2 on every row but the last, the phase equals the
 * unmerged phase of pixel (x - ALIGN16(w)/2, y + 1). */
void oracle_sobel_phase_merged(int w, int h, uint8_t *phase, int phase_stride, const uint8_t *src, int src_stride)
{
    if (h < 3) return;
    const int H = h - 2, half = ALIGN16_C(w) / 2;
    const size_t cells = (size_t)(H - 1) * half + w;
    int16_t *gx = (int16_t *)calloc(cells, sizeof(int16_t)), *gy = (int16_t *)calloc(cells, sizeof(int16_t));
    const size_t rs = (size_t)w * 2;
    int16_t *rx = (int16_t *)malloc(rs * h), *ry = (int16_t *)malloc(rs * h);
    oracle_sobel3x3(w, h, rx, (int)rs, ry, (int)rs, src, src_stride);
    for (int r = 0; r < H; r++)
        for (int x = 0; x < w; x++) {
            gx[(size_t)r * half + x] = rx[(size_t)(r + 1) * w + x];
            gy[(size_t)r * half + x] = ry[(size_t)(r + 1) * w + x];
        }
    for (int r = 0; r < H; r++)
        for (int x = 0; x < w; x++)
            phase[(long)(r + 1) * phase_stride + x] = oracle_phase_px(gx[(size_t)r * half + x], gy[(size_t)r * half + x]);
    free(gx); free(gy); free(rx); free(ry);
}
