/*
 * vision.c -- CPU restatement of the reference Canny (3x3, L1) and Harris (gradient 3, block 3)
 * kernels (TEST INFRASTRUCTURE: the checker, never the product).
 * Memory model: see oracle.h.
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

static inline int rd8(const uint8_t *p, long n, long i) { return (i < 0 || i >= n) ? 0 : p[i]; }

/* HafCpu_FastAtan2_Canny: ago_haf_cpu_arithmetic.cpp:9638-9656 (the statement order matters:
 * gx = gy = 0 ends in sector 2). */
static int canny_sector(int gx, int gy)
{
    unsigned ax = (unsigned)abs(gx) & 0xffff, ay = (unsigned)abs(gy) & 0xffff;
    float d1 = (float)ax * 0.4142135623730950488016887242097f;
    float d2 = (float)ax * 2.4142135623730950488016887242097f;
    int ret = (gx * gy) < 0 ? 3 : 1;
    if ((float)ay <= d1) ret = 0;
    if ((float)ay >= d2) ret = 2;
    return ret;
}

static inline uint16_t canny_pack(int gx, int gy_neg)
{
    int16_t t = (int16_t)(abs((int16_t)gx) + abs((int16_t)gy_neg));
    t = (int16_t)(t << 2);
    t = (int16_t)(t | (canny_sector((int16_t)gx, (int16_t)gy_neg) & 3));
    return (uint16_t)t;
}

/* CannySobel_U16_U8_3x3_L1NORM: ago_haf_cpu_canny.cpp:44-136 (rows 1..h-2; rows 0 and h-1 of the
 * U16 image are never written).  Per pixel: Sobel gx, gy' = -gy, value = ((|gx|+|gy'|) << 2) | sector.
 * Destination rows are 16-byte aligned (prefix 0).  Column partition per row:
 *   - aligned = w - (w & 15); when aligned >= 16 the SSE passes cover [0, aligned): they stage the
 *     vertical sums t0 = p(y-1)+2p(y)+p(y+1) and t1 = p(y+1)-p(y-1) in two scratch rows r0, r1
 *     (r1 = r0 + ALIGN16(w)) and read r0[x-1], r0[x+1], r1[x-1..x+1].  Inside [1, aligned-2] that is
 *     the plain Sobel; at x = 0 it reads r0[-1] (never written: 0) and r1[-1] = r0[ALIGN16(w)-1],
 *     which is this row's t0(w-1) when w is a multiple of 16 and an unwritten 0 otherwise.
 *   - the scalar loop then covers [aligned-1, w) with linear-address reads (it starts at -1 when
 *     aligned == 0, writing the element before the row, i.e. the previous row's last stride slot).
 * Pinned by: bit-exact against oracle/_ref at widths that are / are not multiples of 16 and at
 * widths below 16 (this is where the column-0 and x = -1 behaviours live). */
void oracle_canny_sobel_3x3_l1(int w, int h, uint16_t *dst, int dst_stride, const uint8_t *src, int src_stride)
{
    const long n = (long)h * src_stride;
    const int aligned = w - (w & 15);
    const int a16 = (w + 15) & ~15;
    uint8_t *db = (uint8_t *)dst;
    for (int y = 1; y < h - 1; y++) {
        const long bm = (long)(y - 1) * src_stride, b0 = (long)y * src_stride, bp = (long)(y + 1) * src_stride;
        uint16_t *drow = (uint16_t *)(db + (long)y * dst_stride);
        if (aligned >= 16) {
            for (int x = 0; x < aligned; x++) {
                int gx, gy;
                if (x == 0) {
                    int t0_1 = src[bm + 1] + src[bp + 1] + 2 * src[b0 + 1];
                    int t1_0 = src[bp] - src[bm];
                    int t1_1 = src[bp + 1] - src[bm + 1];
                    int t1_m1 = (w % 16 == 0) ? (src[bm + w - 1] + src[bp + w - 1] + 2 * src[b0 + w - 1]) : 0;
                    gx = t0_1 - 0;
                    gy = -(t1_m1 + 2 * t1_0 + t1_1);
                } else {
                    int t0_m = src[bm + x - 1] + src[bp + x - 1] + 2 * src[b0 + x - 1];
                    int t0_p = (x + 1 < aligned) ? src[bm + x + 1] + src[bp + x + 1] + 2 * src[b0 + x + 1]
                                                 : ((x + 1 == a16) ? (src[bp] - src[bm]) : 0);
                    int t1_m = src[bp + x - 1] - src[bm + x - 1];
                    int t1_0 = src[bp + x] - src[bm + x];
                    int t1_p = (x + 1 < aligned) ? src[bp + x + 1] - src[bm + x + 1] : 0;
                    gx = t0_p - t0_m;
                    gy = -(t1_m + 2 * t1_0 + t1_p);
                }
                drow[x] = canny_pack(gx, gy);
            }
        }
        for (int x = aligned - 1; x < w; x++) {
            long c0 = bm + x, c1 = b0 + x, c2 = bp + x;
            int gx = (rd8(src, n, c0 + 1) - rd8(src, n, c0 - 1)) + (rd8(src, n, c2 + 1) - rd8(src, n, c2 - 1))
                   + 2 * (rd8(src, n, c1 + 1) - rd8(src, n, c1 - 1));
            int gy = rd8(src, n, c2 - 1) + rd8(src, n, c2 + 1) - rd8(src, n, c0 - 1) - rd8(src, n, c0 + 1)
                   + 2 * (rd8(src, n, c2) - rd8(src, n, c0));
            int16_t gy16 = (int16_t)gy;
            gy16 = (int16_t)(~gy16 + 1);
            /* x = -1 writes the 2 bytes before the row (in the previous row's stride) */
            *(uint16_t *)(db + (long)y * dst_stride + 2L * x) = canny_pack((int16_t)gx, gy16);
        }
    }
}

/* CannySuppThreshold_U8XY_U16_3x3: ago_haf_cpu_canny.cpp:549-596, over x in [1, w-2], y in
 * [1, h-2]; neighbours by sector {(-1,0),(1,0)}, {(1,-1),(-1,1)}, {(0,-1),(0,1)}, {(-1,-1),(1,1)}
 * (:26-31); strict '>' against both; out 255 (and pushed) if > upper, 0 if <= lower, else 127.
 * Returns the number of strong pixels (the stack top).
 * Pinned by: bit-exact against oracle/_ref; hand-built 5x5 cases per sector. */
int oracle_canny_supp_threshold(int w, int h, uint8_t *dst, int dst_stride, const uint16_t *mag, int mag_stride,
                                int lower, int upper)
{
    static const int off[4][2][2] = {{{-1, 0}, {1, 0}}, {{1, -1}, {-1, 1}}, {{0, -1}, {0, 1}}, {{-1, -1}, {1, 1}}};
    const int ms = mag_stride / 2;
    int strong = 0;
    for (int y = 1; y < h - 1; y++) {
        for (int x = 1; x < w - 1; x++) {
            const uint16_t *p = mag + (long)y * ms + x;
            int m = p[0] >> 2, a = p[0] & 3;
            int n0 = p[off[a][0][1] * ms + off[a][0][0]] >> 2;
            int n1 = p[off[a][1][1] * ms + off[a][1][0]] >> 2;
            int edge = (m > n0 && m > n1) ? m : 0;
            uint8_t v = edge > upper ? 255 : edge <= lower ? 0 : 127;
            if (v == 255) strong++;
            dst[(long)y * dst_stride + x] = v;
        }
    }
    return strong;
}

/* CannyEdgeTrace_U8_U8XY: ago_haf_cpu_canny.cpp:598-644.  The reference pops a LIFO stack and
 * promotes 8-neighbours equal to 127; the final image is order independent: a pixel is 255 iff it is
 * strong or a weak (127) pixel 8-connected through weak pixels to a strong one; everything else is
 * 0.  Restated as a breadth-first flood from every strong pixel.
 * Pinned by: bit-exact against oracle/_ref; invariant (union-find components) in tests. */
void oracle_canny_trace(int w, int h, uint8_t *img, int stride)
{
    long cap = (long)w * h;
    int *q = (int *)malloc(sizeof(int) * 2 * (cap > 0 ? cap : 1));
    long head = 0, tail = 0;
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++)
            if (img[(long)y * stride + x] == 255) { q[2 * tail] = x; q[2 * tail + 1] = y; tail++; }
    while (head < tail) {
        int x = q[2 * head], y = q[2 * head + 1];
        head++;
        for (int dy = -1; dy <= 1; dy++)
            for (int dx = -1; dx <= 1; dx++) {
                int xx = x + dx, yy = y + dy;
                if ((dx || dy) && xx >= 0 && yy >= 0 && xx < w && yy < h && img[(long)yy * stride + xx] == 127) {
                    img[(long)yy * stride + xx] = 255;
                    q[2 * tail] = xx; q[2 * tail + 1] = yy; tail++;
                }
            }
    }
    free(q);
    int a16 = (w + 15) & ~15;
    if (a16 > stride) a16 = stride;
    for (int y = 0; y < h; y++)
        for (int x = 0; x < a16; x++)
            if (img[(long)y * stride + x] == 127) img[(long)y * stride + x] = 0;
}

/* Layout of the reference's F32x3 HG3 plane (stride S = ALIGN16(12w) bytes).  Both Harris kernels
 * step rows by P = S / sizeof(ago_harris_Gxy_t) = floor(S/12) whole pixels, but the Sobel writer
 * starts its row 1 at byte S (ago_haf_cpu_harris.cpp:96, :166) while the score reader puts row y at
 * y*P pixels (:610, :612).  With S == 12w (w a multiple of 4) both agree; otherwise the reader sees
 * the written floats shifted by (S - 12P)/4 floats.  These helpers give the float offsets. */
static long hg3_write_off(int y, int x, int stride) { long P = stride / 12; return stride / 4 + (long)(y - 1) * 3 * P + 3L * x; }
static long hg3_read_off(int y, int x, int stride) { long P = stride / 12; return (long)y * 3 * P + 3L * x; }

/* HarrisSobel_HG3_U8_3x3: ago_haf_cpu_harris.cpp:85-169.  Rows 1..h-2, all columns, linear reads:
 *   gx = dx(y-1) + 2 dx(y) + dx(y+1), dx(r) = p(x+1,r) - p(x-1,r)
 *   gy = s(y+1) - s(y-1),             s(r)  = p(x-1,r) + 2p(x,r) + p(x+1,r)
 *   out = {(float)gx*gx, (float)gx*gy, (float)gy*gy}   (exact in float for U8 input)
 * stored at the writer offsets above.
 * Pinned by: bit-exact against oracle/_ref (through the whole Harris node, at widths with
 * 12w % 16 == 0 and != 0); equals the products of the Sobel3x3 oracle. */
void oracle_harris_sobel_hg3(int w, int h, float *dst, int dst_stride, const uint8_t *src, int src_stride)
{
    const long n = (long)h * src_stride;
    for (int y = 1; y < h - 1; y++) {
        for (int x = 0; x < w; x++) {
            long bm = (long)(y - 1) * src_stride + x, b0 = (long)y * src_stride + x, bp = (long)(y + 1) * src_stride + x;
            int16_t gx = (int16_t)((rd8(src, n, bp + 1) - rd8(src, n, bp - 1)) + (rd8(src, n, bm + 1) - rd8(src, n, bm - 1))
                                   + ((rd8(src, n, b0 + 1) - rd8(src, n, b0 - 1)) << 1));
            int16_t gy = (int16_t)((rd8(src, n, bp - 1) + (rd8(src, n, bp) << 1) + rd8(src, n, bp + 1))
                                   - (rd8(src, n, bm - 1) + (rd8(src, n, bm) << 1) + rd8(src, n, bm + 1)));
            float *d = dst + hg3_write_off(y, x, dst_stride);
            d[0] = (float)gx * (float)gx;
            d[1] = (float)gx * (float)gy;
            d[2] = (float)gy * (float)gy;
        }
    }
}

/* HarrisScore_HVC_HG3_3x3: ago_haf_cpu_harris.cpp:596-659, norm factor ago_kernel_api.cpp:16941-16942.
 * For y in [1, h-2], x in [1, w-2]: window sums over j = -1..1 (outer), i = -1..1 (inner) from 0,
 * reading HG3 at the reader offsets above;
 *   tr = gx2 + gy2; det = gx2*gy2 - gxy2*gxy2; Mc = (det - (k*tr)*tr) / norm; Vc = Mc > T ? Mc : 0.
 * Rows 0 and h-1 and columns 0 and w-1 are 0.
 * Pinned by: bit-exact against oracle/_ref (the float operation order is the reference's). */
void oracle_harris_score(int w, int h, float *dst, int dst_stride, const float *hg3, int hg3_stride,
                         float k, float thr, float norm)
{
    for (int y = 0; y < h; y++) {
        float *d = (float *)((uint8_t *)dst + (long)y * dst_stride);
        for (int x = 0; x < w; x++) d[x] = 0.0f;
        if (y == 0 || y == h - 1) continue;
        for (int x = 1; x < w - 1; x++) {
            float gx2 = 0, gy2 = 0, gxy2 = 0;
            for (int j = -1; j <= 1; j++) {
                for (int i = -1; i <= 1; i++) {
                    const float *s = hg3 + hg3_read_off(y + j, x + i, hg3_stride);
                    gx2 += s[0];
                    gxy2 += s[1];
                    gy2 += s[2];
                }
            }
            float tr = gx2 + gy2;
            float det = (gx2 * gy2) - (gxy2 * gxy2);
            float mc = det - (k * tr * tr);
            mc /= norm;
            d[x] = (mc > thr) ? mc : 0.0f;
        }
    }
}

/* normFactor of ago_kernel_api.cpp:16941-16942: (255 * 2^(g-1) * 3)^4 in float, left to right */
float oracle_harris_norm(int gradient_size)
{
    float f = 255.0f * (float)(1 << (gradient_size - 1)) * 3;
    return f * f * f * f;
}

/* NonMaxSupp_XY_ANY_3x3: ago_haf_cpu_harris.cpp:882-921; raster order, x in [1, w-2], y in [1, h-2]:
 * keep if >= the three upper neighbours and the left one, and > the right one and the three lower.
 * Stops at capacity.  Items are {int16 x, int16 y, float s}. */
int oracle_nonmax_xys(int w, int h, const float *img, int stride, oracle_xys_t *list, int capacity)
{
    int count = 0;
    for (int y = 1; y < h - 1 && count < capacity; y++) {
        const float *p9 = (const float *)((const uint8_t *)img + (long)(y - 1) * stride);
        const float *p0 = (const float *)((const uint8_t *)img + (long)y * stride);
        const float *p1 = (const float *)((const uint8_t *)img + (long)(y + 1) * stride);
        for (int x = 1; x < w - 1; x++) {
            float c = p0[x];
            if (c >= p9[x - 1] && c >= p9[x] && c >= p9[x + 1] && c >= p0[x - 1] && c > p0[x + 1] &&
                c > p1[x - 1] && c > p1[x] && c > p1[x + 1]) {
                list[count].x = (int16_t)x;
                list[count].y = (int16_t)y;
                list[count].s = c;
                if (++count >= capacity) break;
            }
        }
    }
    return count;
}

static int cmp_key_desc(const void *a, const void *b)
{
    int64_t ka, kb;
    memcpy(&ka, a, 8);
    memcpy(&kb, b, 8);
    return (ka < kb) - (ka > kb);
}

/* HarrisMergeSortAndPick_XY_XYS: ago_haf_cpu_harris.cpp:923-1005, grid set-up ago_kernel_api.cpp:20808-20822.
 * Sort the XYS items as int64 keys (float bits << 32 | y << 16 | x on little-endian) descending;
 * then, when min_distance > 2, greedy pick with a grid of cellSize = floor(d / sqrt 2): reject if
 * the own cell is occupied, else reject if an accepted point in the +-2 cells has
 * dx^2 + dy^2 < ceil(d^2); accepted points fill the output up to capacity and the return value is
 * the uncapped count.  With min_distance <= 2 the first min(n, capacity) sorted items are copied.
 * Sorts `list` in place.
 * Pinned by: bit-exact against oracle/_ref; invariants (sorted output, pairwise distance, maximality). */
int oracle_harris_pick(oracle_xys_t *list, int n, float min_distance, int width, int height,
                       oracle_keypoint_t *out, int capacity)
{
    qsort(list, (size_t)n, sizeof(oracle_xys_t), cmp_key_desc);
    int count = 0;
    if (min_distance > 2.0f) {
        int cell = (int)floor(min_distance / 1.41421356237309504880);
        int gw = (width + cell - 1) / cell, gh = (height + cell - 1) / cell;
        int16_t *grid = (int16_t *)calloc(2 * (size_t)gw * gh, sizeof(int16_t));
        /* HafCpu_MemSet_U32 (ago_haf_cpu_ch_extract_combine.cpp:648-661) stores whole 16-byte vectors,
         * so only the first (cells & ~3) cells become (-1,-1); the last cells % 4 keep the zeroed
         * (0,0) of the node's calloc'd scratch and count as occupied. */
        long cells = (long)gw * gh;
        for (long i = 0; i < 2 * (cells & ~3L); i++) grid[i] = -1;
        int md2 = (int)ceilf(min_distance * min_distance);
        for (int i = 0; i < n; i++) {
            int x = list[i].x, y = list[i].y;
            int cx = x / cell, cy = y / cell;
            int16_t *cg = grid + 2L * (cy * gw + cx);
            int found = 1;
            if (cg[0] < 0) {
                int x0 = cx - 2 < 0 ? 0 : cx - 2, x1 = cx + 2 > gw - 1 ? gw - 1 : cx + 2;
                int y0 = cy - 2 < 0 ? 0 : cy - 2, y1 = cy + 2 > gh - 1 ? gh - 1 : cy + 2;
                found = 0;
                for (int yy = y0; yy <= y1 && !found; yy++)
                    for (int xx = x0; xx <= x1; xx++) {
                        const int16_t *g = grid + 2L * (yy * gw + xx);
                        if (g[0] >= 0) {
                            int dx = g[0] - x, dy = g[1] - y;
                            if (dx * dx + dy * dy < md2) { found = 1; break; }
                        }
                    }
            }
            if (!found) {
                if (count < capacity) {
                    out[count].x = x; out[count].y = y; out[count].strength = list[i].s;
                    out[count].scale = 0.0f; out[count].orientation = 0.0f;
                    out[count].tracking_status = 1; out[count].error = 0.0f;
                }
                count++;
                cg[0] = (int16_t)x;
                cg[1] = (int16_t)y;
            }
        }
        free(grid);
    } else {
        count = n < capacity ? n : capacity;
        for (int i = 0; i < count; i++) {
            out[i].x = list[i].x; out[i].y = list[i].y; out[i].strength = list[i].s;
            out[i].scale = 0.0f; out[i].orientation = 0.0f; out[i].tracking_status = 1; out[i].error = 0.0f;
        }
    }
    return count;
}
