/*
 * fast.c -- CPU restatement of the reference FastCorners kernels (TEST INFRASTRUCTURE: the checker,
 * never the product).  Memory model: see oracle.h (linear reads, 0 outside the plane).
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

static inline int rdp(const uint8_t *p, long n, long i) { return (i < 0 || i >= n) ? 0 : p[i]; }

/* the 16-pixel Bresenham circle of radius 3, clockwise from the top: generateOffset
 * (ago_haf_cpu_fast_corners.cpp:36-56) and the shuffle table dataFastCornersPixelMask (:25-33) */
static const int CIRC[16][2] = {{0, -3}, {1, -3}, {2, -2}, {3, -1}, {3, 0}, {3, 1}, {2, 2}, {1, 3},
                                {0, 3}, {-1, 3}, {-2, 2}, {-3, 1}, {-3, 0}, {-3, -1}, {-2, -2}, {-1, -3}};

/* isCorner (:90-126): nine consecutive set bits on the circle */
static int nine_run(int mask)
{
    if (!mask) return 0;
    mask |= mask << 16;
    for (int i = 0; i < 16; i++, mask >>= 1)
        if ((mask & 0x1ff) == 0x1ff) return 1;
    return 0;
}

/* the early exit shared by every form: pixels 1 and 9, or 5 and 13, both within t of the candidate */
static int early_exit(int c, const int *b, int t)
{
    return (abs(c - b[0]) < t && abs(c - b[8]) < t) || (abs(c - b[4]) < t && abs(c - b[12]) < t);
}

static int plus_mask(int c, const int *b, int t)
{
    int m = 0;
    for (int i = 0; i < 16; i++) if (b[i] > c + t) m |= 1 << i;
    return m;
}

static int minus_mask(int c, const int *b, int t)
{
    int m = 0;
    for (int i = 0; i < 16; i++) if (b[i] < c - t) m |= 1 << i;
    return m;
}

/* checkForCornerAndGetStrength[_SSE] (:216-339, :341-412): after the early exit, for each sign whose
 * mask has a run of nine, the largest threshold in [t, 254] that still passes isCornerPlus/Minus (which
 * repeat the early exit with that threshold), by bisection between t and 255; 0 for the other sign.
 * Returns 0 (no corner) when both masks are empty, else max of the two strengths (which may be 0). */
static int fast_strength(int c, const int *b, int t, int *strength)
{
    if (early_exit(c, b, t)) return 0;
    const int mp = plus_mask(c, b, t), mn = minus_mask(c, b, t);
    int sp = 0, sn = 0;
    if (nine_run(mp)) {
        int lo = t, hi = 255;
        while (hi - lo > 1) {
            const int mid = (hi + lo) >> 1;
            if (!early_exit(c, b, mid) && nine_run(plus_mask(c, b, mid))) lo = mid; else hi = mid;
        }
        sp = lo;
    }
    if (nine_run(mn)) {
        int lo = t, hi = 255;
        while (hi - lo > 1) {
            const int mid = (hi + lo) >> 1;
            if (!early_exit(c, b, mid) && nine_run(minus_mask(c, b, mid))) lo = mid; else hi = mid;
        }
        sn = lo;
    }
    if (!(mp || mn)) return 0;
    *strength = sp > sn ? sp : sn;
    return 1;
}

static void circle(const uint8_t *src, long n, int stride, int x, int y, int *b)
{
    for (int i = 0; i < 16; i++) b[i] = rdp(src, n, (long)(y + CIRC[i][1]) * stride + x + CIRC[i][0]);
}

static void put_kp(oracle_keypoint_t *out, int capacity, unsigned count, int x, int y, float s)
{
    if ((int)count >= capacity) return;
    out[count].x = x;
    out[count].y = y;
    out[count].strength = s;
    out[count].scale = 0.0f;
    out[count].orientation = 0.0f;
    out[count].tracking_status = 1;
    out[count].error = 0.0f;
}

/* HafCpu_FastCorners_XY_U8_NoSupression (ago_haf_cpu_fast_corners.cpp:414-597) and _Supression
 * (:599-790), t = (short)floorf(threshold).
 *   without suppression: candidates at every x in [0, w) of rows 3..h-4 (linear reads reach into the
 *     neighbouring rows at the left and right ends); a corner when the early exit fails and the
 *     plus (b > c + t) or minus (b < c - t) mask has nine consecutive pixels; strength = threshold;
 *   with suppression: strengths (u8) for x in [3, w-4], y in [3, h-4] into a zeroed w*h map, then in
 *     raster order a corner where the strength is nonzero, >= the upper three and left neighbours and
 *     > the right and lower three; strength = that value.
 * Keypoints in raster order, {x, y, strength, 0, 0, 1, 0}; returns the uncapped count.
 * Pinned by: bit-exact against oracle/_ref (both forms, several thresholds, widths with and without
 * an 8-pixel remainder); closed forms on a synthetic step corner. */
int oracle_fast_corners(int w, int h, const uint8_t *src, int stride, float threshold, int nonmax,
                        oracle_keypoint_t *out, int capacity)
{
    const long n = (long)h * stride;
    const int t = (int)(short)floorf(threshold);
    unsigned count = 0;
    int b[16];
    if (!nonmax) {
        for (int y = 3; y < h - 3; y++)
            for (int x = 0; x < w; x++) {
                const int c = rdp(src, n, (long)y * stride + x);
                circle(src, n, stride, x, y, b);
                if (early_exit(c, b, t)) continue;
                if (nine_run(plus_mask(c, b, t)) || nine_run(minus_mask(c, b, t))) {
                    put_kp(out, capacity, count, x, y, threshold);
                    count++;
                }
            }
        return (int)count;
    }
    uint8_t *map = (uint8_t *)calloc((size_t)w * (size_t)(h > 0 ? h : 1), 1);
    for (int y = 3; y < h - 3; y++)
        for (int x = 3; x < w - 3; x++) {
            const int c = rdp(src, n, (long)y * stride + x);
            circle(src, n, stride, x, y, b);
            int s = 0;
            if (fast_strength(c, b, t, &s)) map[(long)y * w + x] = (uint8_t)s;
        }
    for (int y = 3; y < h - 3; y++)
        for (int x = 3; x < w - 3; x++) {
            const uint8_t *p = map + (long)y * w + x;
            const int c = p[0];
            if (c && c >= p[-w - 1] && c >= p[-w] && c >= p[-w + 1] && c >= p[-1] && c > p[1] && c > p[w - 1] && c > p[w] &&
                c > p[w + 1]) {
                put_kp(out, capacity, count, x, y, (float)c);
                count++;
            }
        }
    free(map);
    return (int)count;
}
