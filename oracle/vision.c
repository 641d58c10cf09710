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

/* 16-bit lane arithmetic of the SSE code: wrapping and saturating forms */
static inline int w16(long v) { return (int16_t)(uint16_t)(unsigned long)v; }
static inline int s16(long v) { return v > 32767 ? 32767 : v < -32768 ? -32768 : (int)v; }
static inline int sll16(int v, int k) { return w16((long)((unsigned)(uint16_t)v << k)); }
static inline int abs16(int v) { return w16(v < 0 ? -(long)v : v); }

/* _mm_sqrt_ps of a float, _mm_cvtps_epi32 (nearest even), _mm_packus_epi32 */
static inline int l2_pack_simd(float f)
{
    float r = sqrtf(f);
    long v = (r > -2147483904.0f && r < 2147483648.0f) ? (long)rintf(r) : (long)INT32_MIN;
    return v < 0 ? 0 : v > 65535 ? 65535 : (int)v;
}

/* (vx_int16)sqrt((Gx*Gx) + (Gy*Gy)) of the scalar loops: int products (32-bit wrap), double sqrt,
 * truncation to 32 bits (0x80000000 for NaN), then to 16 */
static inline int l2_scalar(int gx, int gy)
{
    const int32_t q = (int32_t)(uint32_t)((uint64_t)((int64_t)gx * gx) + (uint64_t)((int64_t)gy * gy));
    const double r = sqrt((double)q);
    const int32_t t = (r == r && r < 2147483648.0) ? (int32_t)r : INT32_MIN;
    return (int16_t)(uint16_t)(uint32_t)t;
}

static inline uint16_t canny_out(int mag16, int sector) { return (uint16_t)(sll16(mag16, 2) | (sector & 3)); }

/* CannySobel_U16_U8_{5x5,7x7}_L1NORM and _{3x3,5x5,7x7}_L2NORM: ago_haf_cpu_canny.cpp:145-427 and
 * :646-1056 (the 3x3 L1 form is oracle_canny_sobel_3x3_l1).  Rows R..h-R-1 (R = gsize/2) are written.
 * Each row runs, as the reference does (destination rows 16-byte aligned, so no prefix):
 *   - over [0, aligned) with aligned = w & ~15: a vertical pass into two scratch rows r0, r1 of int16
 *     (r1 = r0 + ALIGN16(w), 8 zero words before r0, zero words after r1), then a horizontal pass that
 *     reads r0/r1 at x-R..x+R -- so column 0 sees zeros on the left and, when w is a multiple of 16,
 *     the end of one scratch row runs into the start of the other.  Vertical / horizontal filters:
 *       5x5: r0 = [1 4 6 4 1] rows, r1 = [-1 -2 0 2 1] rows; Gx = [-1 -2 0 2 1] of r0 (the last add
 *            saturating), Gy = -[1 4 6 4 1] of r1 (L2: the *6 term added saturating);
 *       7x7: r0 = ([1 6 15 20 15 6 1] rows) >> 2, r1 = ([-1 -4 -5 0 5 4 1] rows) >> 2; Gx, Gy with
 *            saturating adds, Gy negated by a saturating subtract;
 *       3x3 L2: r0 = [1 2 1] rows, r1 = [-1 0 1] rows; Gx = r0[x+1] - r0[x-1], Gy = -[1 2 1] of r1;
 *     value: L1 = |Gx| + |Gy| (16-bit wrap; 7x7: signed min with 0x3FFF); L2 = sqrt of the float sum
 *     of squares (3x3: of the low 16 bits of each square) rounded to nearest, saturated to 16 bits;
 *     stored as (value << 2) | sector(Gx, Gy) in 16 bits;
 *   - then over [aligned-1, w) (from -1 when aligned == 0, writing the slot before the row) the
 *     scalar formulas on linear-address reads, in 16-bit wrapping arithmetic: Gx, Gy of the full
 *     kernels (Gy negated for 5x5 L1 only; the 7x7 forms are not divided by 4; the 5x5 L2 Gy adds
 *     the first row's pixels x-1..x+2 where it means to subtract them), L1 = |Gx| + |Gy|,
 *     L2 = (int16)sqrt(Gx*Gx + Gy*Gy).
 * Pinned by: bit-exact against oracle/_ref for each size and norm at widths that are / are not
 * multiples of 16 and below 16; a constant image gives magnitude 0 in every written pixel. */
void oracle_canny_sobel_ex(int w, int h, uint16_t *dst, int dst_stride, const uint8_t *src, int src_stride, int gsize,
                           int l2)
{
    if (gsize == 3 && !l2) {
        oracle_canny_sobel_3x3_l1(w, h, dst, dst_stride, src, src_stride);
        return;
    }
    static const int smooth5[5] = {1, 4, 6, 4, 1}, deriv5[5] = {-1, -2, 0, 2, 1};
    static const int smooth7[7] = {1, 6, 15, 20, 15, 6, 1}, deriv7[7] = {-1, -4, -5, 0, 5, 4, 1};
    static const int smooth3[3] = {1, 2, 1}, deriv3[3] = {-1, 0, 1};
    const int R = gsize / 2;
    const int *sm = gsize == 3 ? smooth3 : gsize == 5 ? smooth5 : smooth7;
    const int *dv = gsize == 3 ? deriv3 : gsize == 5 ? deriv5 : deriv7;
    const long n = (long)h * src_stride;
    const int aligned = w & ~15, a16 = (w + 15) & ~15;
    int16_t *buf = (int16_t *)calloc((size_t)(8 + 2 * a16 + 64), sizeof(int16_t));
    int16_t *r0 = buf + 8, *r1 = r0 + a16;
    uint8_t *db = (uint8_t *)dst;
    for (int y = R; y < h - R; y++) {
        uint16_t *drow = (uint16_t *)(db + (long)y * dst_stride);
        for (int x = 0; x < aligned; x++) {
            int s[7];
            for (int k = 0; k <= 2 * R; k++) s[k] = src[(long)(y - R + k) * src_stride + x];
            int t0, t1;
            if (gsize == 3) {
                t0 = s[0] + s[2] + 2 * s[1];
                t1 = s[2] - s[0];
            } else if (gsize == 5) {
                t0 = 4 * (s[1] + s[3]) + 6 * s[2] + (s[0] + s[4]);
                t1 = 2 * (s[3] - s[1]) + (s[4] - s[0]);
            } else {
                t0 = (6 * (s[1] + s[5]) + 20 * s[3] + (s[0] + s[6]) + 15 * (s[2] + s[4])) >> 2;
                t1 = (4 * (s[5] - s[1]) + (s[6] - s[0]) + 5 * (s[4] - s[2])) >> 2;
            }
            r0[x] = (int16_t)t0;
            r1[x] = (int16_t)t1;
        }
        for (int x = 0; x < aligned; x++) {
#define A0(d) ((int)r0[x + (d)])
#define A1(d) ((int)r1[x + (d)])
            int gx, gy;
            if (gsize == 3) {
                gx = w16(A0(1) - A0(-1));
                gy = w16(-(long)w16(w16(A1(-1) + A1(1)) + sll16(A1(0), 1)));
            } else if (gsize == 5) {
                gx = s16((long)sll16(w16(A0(1) - A0(-1)), 1) + w16(A0(2) - A0(-2)));
                int t = w16((long)sll16(w16(A1(-1) + A1(1)), 2) + w16(A1(-2) + A1(2)));
                t = l2 ? s16((long)t + w16(A1(0) * 6)) : w16((long)t + w16(A1(0) * 6));
                gy = w16(-(long)t);
            } else {
                int t = sll16(s16((long)A0(2) - A0(-2)), 2);
                const int u = w16((long)s16((long)A0(1) - A0(-1)) * 5);
                t = s16((long)t + s16((long)A0(3) - A0(-3)));
                gx = s16((long)t + u);
                int v = s16((long)w16((long)w16(A1(-2) + A1(2)) * 6) + w16((long)A1(0) * 20));
                const int q = w16((long)w16(A1(-1) + A1(1)) * 15);
                v = s16((long)v + s16((long)A1(-3) + A1(3)));
                v = s16((long)v + q);
                gy = s16(-(long)v);
            }
#undef A0
#undef A1
            const int sector = canny_sector(gx, gy);
            int mag;
            if (!l2) {
                mag = w16((long)abs16(gx) + abs16(gy));
                if (gsize == 7 && mag > 0x3fff) mag = 0x3fff;
            } else if (gsize == 3) {
                mag = l2_pack_simd((float)(uint16_t)w16((long)gx * gx) + (float)(uint16_t)w16((long)gy * gy));
            } else {
                mag = l2_pack_simd((float)((long)gx * gx + (long)gy * gy));
            }
            drow[x] = canny_out(mag, sector);
        }
        for (int x = aligned - 1; x < w; x++) {
            long gxs = 0, gys = 0;
            for (int j = 0; j <= 2 * R; j++) {
                const long b = (long)(y - R + j) * src_stride + x;
                for (int i = 0; i <= 2 * R; i++) {
                    const int p = rd8(src, n, b + i - R);
                    gxs += (long)sm[j] * dv[i] * p;
                    /* 5x5 L2: the first row's term lacks its parentheses, so only its first pixel is
                     * subtracted (ago_haf_cpu_canny.cpp:880) */
                    const int wy = (gsize == 5 && l2 && j == 0 && i > 0) ? sm[i] : dv[j] * sm[i];
                    gys += (long)wy * p;
                }
            }
            int gx = (int16_t)(uint16_t)(unsigned long)gxs, gy = (int16_t)(uint16_t)(unsigned long)gys;
            if (gsize == 5 && !l2) gy = (int16_t)(uint16_t)(~(unsigned)gy + 1u);
            int mag = l2 ? l2_scalar(gx, gy) : (int16_t)(uint16_t)(unsigned)(abs(gx) + abs(gy));
            *(uint16_t *)(db + (long)y * dst_stride + 2L * x) = canny_out(mag, canny_sector(gx, gy));
        }
    }
    free(buf);
}

/* CannyEdgeDetector for any gradient size and norm: CANNY_SOBEL_U16_U8_<g>x<g>_<norm> ->
 * CANNY_SUPP_THRESHOLD_U8XY_U16_3x3 (gradient 7: _7x7, which first divides the threshold object's
 * values by 4 in place, ago_kernel_api.cpp:18033-18035) -> CANNY_EDGE_TRACE_U8_U8XY
 * (ago_drama_divide.cpp:1618-1701).  lower/upper are the values the suppression reads. */
void oracle_canny_node_ex(int w, int h, const uint8_t *src, int gsize, int l2, int lower, int upper, uint8_t *dst)
{
    const int s16w = ((w + 15) & ~15), s8w = ((2 * w + 15) & ~15);
    uint8_t *in = (uint8_t *)calloc((size_t)s16w * h + 64, 1), *mag = (uint8_t *)calloc((size_t)s8w * h + 64, 1);
    uint8_t *out = (uint8_t *)calloc((size_t)s16w * h + 64, 1);
    for (int y = 0; y < h; y++) memcpy(in + (long)y * s16w, src + (long)y * w, (size_t)w);
    oracle_canny_sobel_ex(w, h, (uint16_t *)mag, s8w, in, s16w, gsize, l2);
    oracle_canny_supp_threshold(w, h, out, s16w, (const uint16_t *)mag, s8w, lower, upper);
    oracle_canny_trace(w, h, out, s16w);
    for (int y = 0; y < h; y++) memcpy(dst + (long)y * w, out + (long)y * s16w, (size_t)w);
    free(in); free(mag); free(out);
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

/* HarrisSobel_HG3_U8_5x5 / _7x7: ago_haf_cpu_harris.cpp:333-453 / 455-594.  Per source row r the
 * horizontal passes (linear reads) hd(r) = derivative taps, hs(r) = smoothing taps (5: [-1 -2 0 2 1],
 * [1 4 6 4 1] in int16; 7: [-1 -4 -5 0 5 4 1], [1 6 15 20 15 6 1] in int32); for output rows
 * y = R..h-R-1: gx = smoothing of hd down the rows, gy = derivative of hs down the rows;
 * out = {(float)gx*gx, (float)gx*gy, (float)gy*gy}.  The writer starts at byte R*stride and then
 * advances one pixel per pixel with no row padding (packed rows of w pixels), unlike the 3x3 writer.
 * Pinned by: bit-exact against oracle/_ref through the whole Harris node (gradient 5 and 7 with every
 * block size, at widths with 12w % 16 == 0 and != 0). */
void oracle_harris_sobel_g(int w, int h, float *dst, int dst_stride, const uint8_t *src, int src_stride, int g)
{
    static const int d5[5] = {-1, -2, 0, 2, 1}, s5[5] = {1, 4, 6, 4, 1};
    static const int d7[7] = {-1, -4, -5, 0, 5, 4, 1}, s7[7] = {1, 6, 15, 20, 15, 6, 1};
    const int R = g / 2;
    const int *dv = g == 5 ? d5 : d7, *sm = g == 5 ? s5 : s7;
    const long n = (long)h * src_stride;
    for (int y = R; y < h - R; y++) {
        for (int x = 0; x < w; x++) {
            long gx = 0, gy = 0;
            for (int j = 0; j < g; j++) {
                const long b = (long)(y - R + j) * src_stride + x;
                long hd = 0, hs = 0;
                for (int i = 0; i < g; i++) {
                    const int p = rd8(src, n, b + i - R);
                    hd += (long)dv[i] * p;
                    hs += (long)sm[i] * p;
                }
                gx += (long)sm[j] * hd;
                gy += (long)dv[j] * hs;
            }
            float *d = dst + (long)R * (dst_stride / 4) + 3L * ((long)(y - R) * w + x);
            d[0] = (float)gx * (float)gx;
            d[1] = (float)gx * (float)gy;
            d[2] = (float)gy * (float)gy;
        }
    }
}

/* HarrisScore_HVC_HG3_5x5 / _7x7: ago_haf_cpu_harris.cpp:661-729 / 731-802, reading HG3 at pixel
 * y*P + x (P = floor(stride/12)); window sums j-major over -R..R from 0, then as for 3x3.  5x5: output
 * rows 2..h-3, columns 2..w-3; the rest 0.  7x7: rows 3..h-4, and the value computed at x (3..w-4) is
 * stored at column x-1 (its third leading zero does not advance the pointer), columns 0..1 and
 * w-4..w-2 are 0 and column w-1 is never written (0 in a zeroed plane).
 * Pinned by: bit-exact against oracle/_ref through the whole Harris node. */
void oracle_harris_score_b(int w, int h, float *dst, int dst_stride, const float *hg3, int hg3_stride, float k,
                           float thr, float norm, int b)
{
    const int R = b / 2;
    for (int y = 0; y < h; y++) {
        float *d = (float *)((uint8_t *)dst + (long)y * dst_stride);
        for (int x = 0; x < w; x++) d[x] = 0.0f;
        if (y < R || y >= h - R) continue;
        for (int x = R; x < w - R; x++) {
            float gx2 = 0, gy2 = 0, gxy2 = 0;
            for (int j = -R; j <= R; j++) {
                for (int i = -R; i <= R; i++) {
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
            d[b == 7 ? x - 1 : x] = (mc > thr) ? mc : 0.0f;
        }
    }
}

/* normFactor of ago_kernel_api.cpp:16941-16942, :17012-17013, :17083-17084: (255 * 2^(g-1) * b)^4 */
float oracle_harris_norm_gb(int gradient_size, int block_size)
{
    float f = 255.0f * (float)(1 << (gradient_size - 1)) * (float)block_size;
    return f * f * f * f;
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
