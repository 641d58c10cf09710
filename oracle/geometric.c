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
