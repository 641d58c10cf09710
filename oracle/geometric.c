/*
 * geometric.c -- CPU restatement of ColorConvert luma, ScaleImage (bilinear) and the HALF
 * Gaussian pyramid (TEST INFRASTRUCTURE: the checker, never the product).
 * Memory model: see oracle.h (linear addressing with the zero guard, reference strides).
 */
#include "oracle.h"
#include <math.h>
#include <string.h>

static inline int rd8(const uint8_t *p, long n, long i) { return (i < 0 || i >= n) ? 0 : p[i]; }

/* ColorConvert_Y_RGB: ago_haf_cpu_color_convert.cpp:2913-2990.  Per pixel, in float:
 * ((R * 0.2126f) + G * 0.7152f) + B * 0.0722f, truncated (cvttps) and saturated to U8 (packus).
 * The scalar tail (:2979-2986) computes the same expression with a (vx_uint8) cast.
 * Pinned by: bit-exact against oracle/_ref; closed forms for grey pixels (R=G=B=v gives v or
 * v-1 depending on the float sum of the weights, checked exhaustively) and pure primaries. */
void oracle_color_y_rgb(int w, int h, uint8_t *dst, int dst_stride, const uint8_t *rgb, int rgb_stride)
{
    const float wr = (float)0.2126, wg = (float)0.7152, wb = (float)0.0722;
    for (int y = 0; y < h; y++) {
        const uint8_t *s = rgb + (long)y * rgb_stride;
        for (int x = 0; x < w; x++) {
            float v = (float)s[3 * x] * wr;
            v = v + (float)s[3 * x + 1] * wg;
            v = v + (float)s[3 * x + 2] * wb;
            int iv = (int)v;
            dst[(long)y * dst_stride + x] = (uint8_t)(iv < 0 ? 0 : iv > 255 ? 255 : iv);
        }
    }
}

/* ScaleImage_U8_U8_Bilinear: ago_haf_cpu_geometric.cpp:1977-2105 with the scale matrix of
 * ago_kernel_api.cpp:19897-19916.
 *   xscale = (f32)(srcW/dstW), xoffset = (f32)(srcW/dstW*0.5 - 0.5) (double math, then float);
 *   xinc = (int)(2^18 * xscale), xoffs = (int)(2^18 * xoffset) (truncation toward zero);
 *   xpos = xoffs + x*xinc, xmap = max(xpos >> 18, 0) (the right clamp at :2009-2012 is dead),
 *   xf = ((xpos & 0x3ffff) + 0x200) >> 10; rows alike, with ym < 0 -> both rows 0 and
 *   ym >= srcH-1 -> both rows srcH-1.
 *   Body columns x < (dstW & ~7): two roundings, t = ((256-xf) p0 + xf p1 + 128) >> 8 per row then
 *   ((256-yf) t0 + yf t1 + 128) >> 8, in 16-bit lanes; tail columns: one rounding
 *   (sum of the four weighted pixels + 0x8000) >> 16.
 *   p1 = the byte after p0 (a linear read, so xmap = srcW-1 reads the next row / padding).
 * Pinned by: bit-exact against oracle/_ref at 2:1, 1:2 and odd ratios; closed form at exactly 2:1
 * (xf = yf = 128: two-stage rounded averages of 2x2 blocks). */
void oracle_scale_bilinear(int dw, int dh, uint8_t *dst, int dst_stride, int sw, int sh, const uint8_t *src,
                           int src_stride)
{
    const long n = (long)sh * src_stride;
    const float xscale = (float)((double)sw / (double)dw);
    const float yscale = (float)((double)sh / (double)dh);
    const float xoffset = (float)((double)sw / (double)dw * 0.5 - 0.5);
    const float yoffset = (float)((double)sh / (double)dh * 0.5 - 0.5);
    const int xinc = (int)((float)(1 << 18) * xscale), yinc = (int)((float)(1 << 18) * yscale);
    const int xoffs = (int)((float)(1 << 18) * xoffset), yoffs = (int)((float)(1 << 18) * yoffset);
    const int body = dw & ~7;
    for (int y = 0, ypos = yoffs; y < dh; y++, ypos += yinc) {
        int ym = ypos >> 18;
        int yf = ((ypos & 0x3ffff) + 0x200) >> 10;
        int omy = 0x100 - yf;
        long r0, r1;
        if (ym < 0) { r0 = r1 = 0; }
        else if (ym >= sh - 1) { r0 = r1 = (long)(sh - 1) * src_stride; }
        else { r0 = (long)ym * src_stride; r1 = r0 + src_stride; }
        for (int x = 0; x < dw; x++) {
            int xpos = xoffs + x * xinc;
            int xm = xpos >> 18;
            if (xm < 0) xm = 0;
            int xf = ((xpos & 0x3ffff) + 0x200) >> 10;
            int omx = 0x100 - xf;
            int p00 = rd8(src, n, r0 + xm), p01 = rd8(src, n, r0 + xm + 1);
            int p10 = rd8(src, n, r1 + xm), p11 = rd8(src, n, r1 + xm + 1);
            int v;
            if (x < body) {
                int t0 = ((omx * p00 + xf * p01 + 0x80) & 0xffff) >> 8;
                int t1 = ((omx * p10 + xf * p11 + 0x80) & 0xffff) >> 8;
                v = ((omy * t0 + yf * t1 + 0x80) & 0xffff) >> 8;
                if (v > 255) v = 255;
            } else {
                v = (omx * omy * p00 + xf * omy * p01 + omx * yf * p10 + xf * yf * p11 + 0x8000) >> 16;
            }
            dst[(long)y * dst_stride + x] = (uint8_t)v;
        }
    }
}

/* The scale matrix of ago_kernel_api.cpp:19757-19800 (nearest: offset = ratio/2), :19897-19916
 * (bilinear forms: offset = ratio/2 - 1/2) and :20242-20275 (area: offset = -1/2), in float from
 * double arithmetic; the fixed-point forms truncate 2^18 * value toward zero.  Positions advance by
 * repeated 32-bit additions (ago_haf_cpu_geometric.cpp:1897-1918), i.e. modulo 2^32. */
static void scale_matrix(int dw, int dh, int sw, int sh, double half, int *xinc, int *yinc, int *xoffs, int *yoffs,
                         float *xscale, float *yscale)
{
    *xscale = (float)((double)sw / (double)dw);
    *yscale = (float)((double)sh / (double)dh);
    const float xoffset = (float)((double)sw / (double)dw * 0.5 - half);
    const float yoffset = (float)((double)sh / (double)dh * 0.5 - half);
    *xinc = (int)((float)(1 << 18) * *xscale);
    *yinc = (int)((float)(1 << 18) * *yscale);
    *xoffs = (int)((float)(1 << 18) * xoffset);
    *yoffs = (int)((float)(1 << 18) * yoffset);
}

static inline int pos_at(int offs, int inc, int i) { return (int)((uint32_t)offs + (uint32_t)inc * (uint32_t)i); }

/* ScaleImage_U8_U8_Nearest: ago_haf_cpu_geometric.cpp:1874-1976 (offset ratio/2, no -1/2).
 *   map = pos >> 18 clamped to [0, size-1] per axis; out = src[ymap][xmap].
 * Pinned by: bit-exact against oracle/_ref (up- and downscale, odd sizes); integer ratios give the
 * closed form out(x,y) = src(k x + k/2, k y + k/2) for an exact downscale by k. */
void oracle_scale_nearest(int dw, int dh, uint8_t *dst, int dst_stride, int sw, int sh, const uint8_t *src,
                          int src_stride)
{
    int xinc, yinc, xoffs, yoffs;
    float xs, ys;
    scale_matrix(dw, dh, sw, sh, 0.0, &xinc, &yinc, &xoffs, &yoffs, &xs, &ys);
    for (int y = 0; y < dh; y++) {
        int ym = pos_at(yoffs, yinc, y) >> 18;
        if (ym > sh - 1) ym = sh - 1;
        if (ym < 0) ym = 0;
        for (int x = 0; x < dw; x++) {
            int xm = pos_at(xoffs, xinc, x) >> 18;
            if (xm > sw - 1) xm = sw - 1;
            if (xm < 0) xm = 0;
            dst[(long)y * dst_stride + x] = src[(long)ym * src_stride + xm];
        }
    }
}

/* The column split of the bilinear border forms (ago_haf_cpu_geometric.cpp:2133-2155, :2297-2319):
 * newDstWidth = (first x whose map reaches sw-1) - 1, where the value 0 counts as unset (a later x
 * sets it again); 0 when no column reaches sw-1; rounded down to a multiple of 8 when dw is not one.
 * Columns below it take the two-stage (SIMD) rounding, the rest the single rounding. */
int oracle_scale_border_split(int dw, int sw, int xoffs, int xinc)
{
    uint32_t nd = 0;
    for (int x = 0; x < dw; x++)
        if ((pos_at(xoffs, xinc, x) >> 18) >= sw - 1 && !nd) nd = (uint32_t)x - 1u;
    if (dw & 7) nd &= ~7u;
    return (int)nd;
}

/* ScaleImage_U8_U8_Bilinear_Replicate (ago_haf_cpu_geometric.cpp:2108-2258) and _Constant
 * (:2260-2391), scale matrix as the UNDEFINED form.  Per output row: ym = ypos >> 18, yf =
 * ((ypos & 0x3ffff) + 0x200) >> 10; rows (R1, R2) = (ym, ym+1), or for ym < 0 (0, 0) | (border, 0),
 * for ym >= sh-1 (sh-1, sh-1) | (sh-1, border).  Per column: Xmap = xpos >> 18 clamped to [0, sw-1],
 * xf likewise from xpos.
 *   x < split: t1 = ((256-xf) R1[Xmap] + xf R1[Xmap+1] + 0x80) >> 8, t2 from R2, then
 *              ((256-yf) t1 + yf t2 + 0x80) >> 8 (16-bit lanes; no term reaches 2^16);
 *   x >= split: ((256-xf)(256-yf) p0 + xf(256-yf) p1 + (256-xf) yf p2 + xf yf p3 + 0x8000) >> 16 with
 *              p1 (p3) = the right neighbour, or at Xmap = sw-1 p0 (p2) | the border value.
 *   Upscale (xscale < 1) with split > 0: column 0 again with the left neighbour replaced:
 *              replicate p0 = p1 = R1[0], p2 = p3 = R2[0]; constant p0 = p2 = border, p1 = R1[0],
 *              p3 = R2[0] (with split 0 the single-rounding pass rewrites column 0 afterwards).
 * Pinned by: bit-exact against oracle/_ref for up/downscale, sizes on and off multiples of 8, border
 * values 0 and 200; constant image => constant output (replicate), impulse responses. */
void oracle_scale_bilinear_border(int dw, int dh, uint8_t *dst, int dst_stride, int sw, int sh, const uint8_t *src,
                                  int src_stride, int constant, int border)
{
    int xinc, yinc, xoffs, yoffs;
    float xs, ys;
    scale_matrix(dw, dh, sw, sh, 0.5, &xinc, &yinc, &xoffs, &yoffs, &xs, &ys);
    const int split = oracle_scale_border_split(dw, sw, xoffs, xinc);
    for (int y = 0; y < dh; y++) {
        const int ypos = pos_at(yoffs, yinc, y);
        const int ym = ypos >> 18;
        const int yf = ((ypos & 0x3ffff) + 0x200) >> 10, omy = 0x100 - yf;
        int r1, r2;   /* -1: the border row */
        if (ym < 0) { r1 = constant ? -1 : 0; r2 = 0; }
        else if (ym >= sh - 1) { r1 = sh - 1; r2 = constant ? -1 : sh - 1; }
        else { r1 = ym; r2 = ym + 1; }
#define PX(r, c) ((r) < 0 ? border : (int)src[(long)(r) * src_stride + (c)])
        for (int x = 0; x < dw; x++) {
            const int xpos = pos_at(xoffs, xinc, x);
            int xm = xpos >> 18;
            if (xm >= sw - 1) xm = sw - 1;
            else if (xm < 0) xm = 0;
            const int xf = ((xpos & 0x3ffff) + 0x200) >> 10, omx = 0x100 - xf;
            int v;
            if (x < split) {
                const int t1 = (omx * PX(r1, xm) + xf * PX(r1, xm + 1) + 0x80) >> 8;
                const int t2 = (omx * PX(r2, xm) + xf * PX(r2, xm + 1) + 0x80) >> 8;
                v = (omy * t1 + yf * t2 + 0x80) >> 8;
            } else {
                const int p0 = PX(r1, xm), p2 = PX(r2, xm);
                const int p1 = xm < sw - 1 ? PX(r1, xm + 1) : (constant ? border : p0);
                const int p3 = xm < sw - 1 ? PX(r2, xm + 1) : (constant ? border : p2);
                v = (int)(((unsigned)(omx * omy * p0) + (unsigned)(xf * omy * p1) + (unsigned)(omx * yf * p2) +
                           (unsigned)(xf * yf * p3) + 0x8000u) >> 16);
            }
            dst[(long)y * dst_stride + x] = (uint8_t)v;
        }
        if (xs < 1.0f && split > 0) {
            const int xpos = xoffs;
            const int xf = ((xpos & 0x3ffff) + 0x200) >> 10, omx = 0x100 - xf;
            int p0, p1, p2, p3;
            if (constant) { p0 = p2 = border; p1 = PX(r1, 0); p3 = PX(r2, 0); }
            else { p0 = p1 = PX(r1, 0); p2 = p3 = PX(r2, 0); }
            dst[(long)y * dst_stride] = (uint8_t)(((unsigned)(omx * omy * p0) + (unsigned)(xf * omy * p1) +
                                                   (unsigned)(omx * yf * p2) + (unsigned)(xf * yf * p3) + 0x8000u) >> 16);
        }
#undef PX
    }
}

/* ScaleImage_U8_U8_Area as a graph runs it.  The reference's graph optimizer rewrites every AMD
 * ScaleImage node before it runs (ago_drama_remove.cpp:1270-1311): the same size becomes
 * CHANNEL_COPY_U8_U8, and an AREA node that does not shrink both dimensions becomes
 * SCALE_IMAGE_U8_U8_NEAREST.  What reaches HafCpu_ScaleImage_U8_U8_Area (ago_haf_cpu_geometric.cpp:
 * 2491-2617, offsets -1/2 from ago_kernel_api.cpp:20276-20277) has a defined result for an exact 2x
 * only: row y averages source rows S = (int)(2y - 0.5f) clamped to [0, sh-2] and S+1,
 * (a + b + c + d + 2) >> 2 over 2x2 blocks, columns 0..(dw & ~7)-1 only (the rest are not written).
 * Returns -1 for other shrinking ratios: the generic path keeps its row sums past the end of its node
 * buffer (localDataSize = sizeof(matrix) + 2*alignedWidth + 16 ends where Ymap starts), so its result
 * is not defined; also for a 2x destination narrower than 8 (the loop bound dw - 8 wraps).
 * Pinned by: bit-exact against oracle/_ref (the named AMD kernel and vxScaleImageNode) on the same
 * size, upscales, mixed ratios and exact 2x sizes; constant image => constant output. */
int oracle_scale_area(int dw, int dh, uint8_t *dst, int dst_stride, int sw, int sh, const uint8_t *src,
                      int src_stride)
{
    if (sw == dw && sh == dh) {
        for (int y = 0; y < dh; y++) memcpy(dst + (long)y * dst_stride, src + (long)y * src_stride, (size_t)dw);
        return 0;
    }
    if (!(sw > dw && sh > dh)) {
        oracle_scale_nearest(dw, dh, dst, dst_stride, sw, sh, src, src_stride);
        return 0;
    }
    const float xs = (float)((double)sw / (double)dw), ys = (float)((double)sh / (double)dh);
    if (xs == 2.0f && ys == 2.0f && dw >= 8 && sh >= 2) {
        for (int y = 0; y < dh; y++) {
            int r = (int)(-0.5f + (float)(unsigned)(2 * y));
            if (r < 0) r = 0;
            if (r > sh - 2) r = sh - 2;
            const uint8_t *s0 = src + (long)r * src_stride, *s1 = s0 + src_stride;
            for (int x = 0; x < (dw & ~7); x++)
                dst[(long)y * dst_stride + x] = (uint8_t)((s0[2 * x] + s0[2 * x + 1] + s1[2 * x] + s1[2 * x + 1] + 2) >> 2);
        }
        return 0;
    }
    return -1;
}

/* ColorConvert_RGB_IYUV / _RGB_NV12: ago_haf_cpu_color_convert.cpp:472-623 / 624-768 (BT.709).  Each 2x2
 * block shares one chroma sample (IYUV: separate U, V planes; NV12: interleaved U,V); u = U - 128,
 * v = V - 128 in float; per pixel, truncated and clamped to [0, 255]:
 *   R = Y + v*1.5748, G = Y - (u*0.1873 + v*0.4681), B = Y + u*1.8556
 * (the SSE body adds u*0 or v*0 terms, which do not change the sums, and truncates before saturating;
 * the scalar tail clamps before truncating -- the same bytes).  w and h even.
 * Pinned by: bit-exact against oracle/_ref (color_convert IYUV -> RGB, NV12 -> RGB); grey chroma
 * (128) gives R = G = B = Y. */
void oracle_color_rgb_yuv(int w, int h, uint8_t *dst, int dst_stride, const uint8_t *y, int y_stride, const uint8_t *u,
                          int u_stride, const uint8_t *v, int v_stride, int nv12)
{
    for (int r = 0; r < h; r++)
        for (int x = 0; x < w; x++) {
            const float Y = (float)y[(long)r * y_stride + x];
            float U, V;
            if (nv12) {
                U = (float)u[(long)(r / 2) * u_stride + 2 * (x / 2)] - 128.0f;
                V = (float)u[(long)(r / 2) * u_stride + 2 * (x / 2) + 1] - 128.0f;
            } else {
                U = (float)u[(long)(r / 2) * u_stride + x / 2] - 128.0f;
                V = (float)v[(long)(r / 2) * v_stride + x / 2] - 128.0f;
            }
            const float G = (U * 0.1873f) + (V * 0.4681f), R = V * 1.5748f, B = U * 1.8556f;
            const float c[3] = {Y + R, Y - G, Y + B};
            for (int k = 0; k < 3; k++) {
                const float f = fminf(fmaxf(c[k], 0.0f), 255.0f);
                dst[(long)r * dst_stride + 3L * x + k] = (uint8_t)f;
            }
        }
}

/* ScaleGaussianHalf_U8_U8_3x3: ago_haf_cpu_geometric.cpp:2772-2829, wrapper ago_kernel_api.cpp:15341-15352
 * (rows start one row in on both images; zeroed node scratch, ago_interface.cpp:1859-1866).
 * Output rows 1..dh-2, all columns; output row y is centred on source row 2y whatever the parity:
 *   V(c) = p(c, 2y-1) + 2 p(c, 2y) + p(c, 2y+1) for c >= 0, V(-1) = 0 (the zeroed scratch word
 *          before the row buffer), reads inside the source row stride (columns up to sw read 0 padding)
 *   out  = (V(2x-1) + 2 V(2x) + V(2x+1)) >> 4                             (no rounding)
 * Returns -1 when dh < 2 (the reference's dstHeight - 2 wraps) or the output is larger than half the
 * source rounded up.
 * Pinned by: bit-exact against oracle/_ref (halfscale_gaussian with kernel size 3) on even/odd
 * sizes; constant image => constant interior with column 0 at 3/4; impulse => [1 2 1]^2 / 16. */
int oracle_scale_gaussian_half_3x3(int dw, int dh, uint8_t *dst, int dst_stride, int sw, int sh,
                                   const uint8_t *src, int src_stride)
{
    if (dh < 2 || dw > (sw + 1) / 2 || dh > (sh + 1) / 2) return -1;
    for (int y = 1; y < dh - 1; y++) {
        const uint8_t *r0 = src + (long)(2 * y - 1) * src_stride, *r1 = r0 + src_stride, *r2 = r1 + src_stride;
#define V(c) ((c) < 0 ? 0u : (unsigned)r0[c] + 2u * (unsigned)r1[c] + (unsigned)r2[c])
        for (int x = 0; x < dw; x++)
            dst[(long)y * dst_stride + x] = (uint8_t)((V(2 * x - 1) + 2u * V(2 * x) + V(2 * x + 1)) >> 4);
#undef V
    }
    return 0;
}

/* ScaleGaussianHalf_U8_U8_5x5: ago_haf_cpu_pyramid.cpp:94-305, wrapper ago_kernel_api.cpp:15420-15436.
 * Output rows 1..dh-2, all columns:
 *   cy = 2y + oy (oy = 1 if srcH is even), cx = 2x + ox (ox = 1 if srcW is even);
 *   H(r) = p(cx-2,r) + 4p(cx-1,r) + 6p(cx,r) + 4p(cx+1,r) + p(cx+2,r)    (linear reads, u16)
 *   out = (H(cy-2) + 4H(cy-1) + 6H(cy) + 4H(cy+1) + H(cy+2)) >> 8        (u16, no rounding)
 * Returns -1 (VX_ERROR_INVALID_DIMENSION in the reference) when src < 5 or dst < 3 in a dimension.
 * Pinned by: bit-exact against oracle/_ref on even/odd sizes; constant image => constant
 * interior; impulse => the separable [1 4 6 4 1]^2 / 256 kernel. */
int oracle_scale_gaussian_half_5x5(int dw, int dh, uint8_t *dst, int dst_stride, int sw, int sh,
                                   const uint8_t *src, int src_stride)
{
    if (sw < 5 || sh < 5 || dw < 3 || dh < 3) return -1;
    const long n = (long)sh * src_stride;
    const int oy = (sh & 1) ? 0 : 1, ox = (sw & 1) ? 0 : 1;
    static const int k[5] = {1, 4, 6, 4, 1};
    for (int y = 1; y < dh - 1; y++) {
        int cy = 2 * y + oy;
        for (int x = 0; x < dw; x++) {
            int cx = 2 * x + ox;
            unsigned acc = 0;
            for (int j = 0; j < 5; j++) {
                long b = (long)(cy + j - 2) * src_stride + cx;
                unsigned hr = 0;
                for (int i = 0; i < 5; i++) hr += (unsigned)k[i] * (unsigned)rd8(src, n, b + i - 2);
                acc += (unsigned)k[j] * hr;
            }
            dst[(long)y * dst_stride + x] = (uint8_t)((acc & 0xffff) >> 8);
        }
    }
    return 0;
}

/* ColorConvert RGB -> IYUV: HafCpu_ColorConvert_IYUV_RGB (ago_haf_cpu_color_convert.cpp:2492-2705), per
 * 2x2 block.  Columns below w & ~3 follow the SSE body: per pixel, in float, Y = ((R*0.2126f) + G*0.7152f)
 * + B*0.0722f, U = (((R*-0.1146f) + G*-0.3854f) + B*0.5f) + 128, V = (((R*0.5f) + G*-0.4542f) +
 * B*-0.0458f) + 128, each truncated; chroma = ((u00+u10+1)>>1 + (u01+u11+1)>>1 + 1) >> 1 (a vertical
 * rounding average, then a horizontal one).  The last two columns when w % 4 == 2 follow the scalar tail:
 * Y adds B*0.0722 in double, and U / V are the float sum of the four pixel values (row-major order)
 * divided by 4, truncated. */
static float yuv_u_px(float r, float g, float b) { return ((r * -0.1146f) + (g * -0.3854f)) + (b * 0.5f) + 128.0f; }
static float yuv_v_px(float r, float g, float b) { return ((r * 0.5f) + (g * -0.4542f)) + (b * -0.0458f) + 128.0f; }
static int sat_trunc(float f, int hi) { int i = (int)f; return i < 0 ? 0 : (i > hi ? hi : i); }

void oracle_color_iyuv_from_rgb(int w, int h, const uint8_t *rgb, int rgbStride, uint8_t *y, int yStride, uint8_t *u,
                                int uStride, uint8_t *v, int vStride)
{
    const int aligned = w & ~3;
    for (int r = 0; r < h; r += 2)
        for (int x = 0; x < w; x += 2) {
            float R[4], G[4], B[4];
            for (int k = 0; k < 4; k++) {
                const uint8_t *p = rgb + (long)(r + (k >> 1)) * rgbStride + 3L * (x + (k & 1));
                R[k] = p[0]; G[k] = p[1]; B[k] = p[2];
            }
            int uo, vo;
            if (x < aligned) {
                int uu[4], vv[4];
                for (int k = 0; k < 4; k++) {
                    const float yf = ((R[k] * 0.2126f) + (G[k] * 0.7152f)) + (B[k] * 0.0722f);
                    y[(long)(r + (k >> 1)) * yStride + x + (k & 1)] = (uint8_t)sat_trunc(yf, 255);
                    uu[k] = sat_trunc(yuv_u_px(R[k], G[k], B[k]), 65535);
                    vv[k] = sat_trunc(yuv_v_px(R[k], G[k], B[k]), 65535);
                }
                uo = ((((uu[0] + uu[2] + 1) >> 1) + ((uu[1] + uu[3] + 1) >> 1)) + 1) >> 1;
                vo = ((((vv[0] + vv[2] + 1) >> 1) + ((vv[1] + vv[3] + 1) >> 1)) + 1) >> 1;
                if (uo > 255) uo = 255;
                if (vo > 255) vo = 255;
            } else {
                float U = 0.0f, V = 0.0f;
                for (int k = 0; k < 4; k++) {
                    const double yd = (double)((R[k] * 0.2126f) + (G[k] * 0.7152f)) + (double)B[k] * 0.0722;
                    y[(long)(r + (k >> 1)) * yStride + x + (k & 1)] = (uint8_t)yd;
                    U = k ? U + yuv_u_px(R[k], G[k], B[k]) : yuv_u_px(R[k], G[k], B[k]);
                    V = k ? V + yuv_v_px(R[k], G[k], B[k]) : yuv_v_px(R[k], G[k], B[k]);
                }
                U /= 4.0f;
                V /= 4.0f;
                uo = (uint8_t)U;
                vo = (uint8_t)V;
            }
            u[(long)(r >> 1) * uStride + (x >> 1)] = (uint8_t)uo;
            v[(long)(r >> 1) * vStride + (x >> 1)] = (uint8_t)vo;
        }
}
