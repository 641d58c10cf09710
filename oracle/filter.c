/*
 * filter.c -- CPU restatement of the reference 3x3 U8 filters (TEST INFRASTRUCTURE: the checker,
 * never the product).  Memory model: see oracle.h.
 */
#include "oracle.h"
#include <string.h>

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
