// geometric.hip -- ColorConvert luma (RGB -> Y), ChannelCopy, ScaleImage bilinear and the HALF
// Gaussian pyramid step (ScaleGaussianHalf 5x5) for gfx950.
//
// All four are HBM-bound streaming kernels.  Reads that the reference performs at linear addresses
// (the pyramid's x-2..x+2 taps, the bilinear kernel's p(xmap+1)) go through 16-byte chunk loads that
// return zeros for chunks outside [0, h*stride) -- the zero guard of the reference layout.
#include "mvx_common.hpp"
#include "pyramid.hpp"

namespace mvx {

inline bool plane_ok_g(const void *p, uint32_t w, uint32_t stride, uint32_t bpp)
{
    return p && aligned16(p) && (stride % 16u) == 0 && stride >= ((w * bpp + 15u) & ~15u);
}

// ------------------------------------------------------------------ ColorConvert Y <- RGB ------
// HafCpu_ColorConvert_Y_RGB (ago_haf_cpu_color_convert.cpp:2913-2990): per pixel, in float,
// ((R * 0.2126f) + G * 0.7152f) + B * 0.0722f, truncated, saturated to U8.
__global__ void __launch_bounds__(256)
k_color_y_rgb(uint32_t w, uint32_t h, uint8_t *__restrict__ dst, uint32_t dstStride, const uint8_t *__restrict__ rgb,
              uint32_t rgbStride)
{
    const uint32_t x0 = (blockIdx.x * 256 + threadIdx.x) * 16;
    const uint32_t y = blockIdx.y;
    if (x0 >= w || y >= h) return;
    const uint8_t *s = rgb + (long)y * rgbStride + 3L * x0;
    const int nv = min(16, (int)(w - x0));
    uint32_t px[48];
    if (nv == 16) {
        uint32_t a[16], b[16], c[16];
        unpack16(*reinterpret_cast<const uint4 *>(s), a);
        unpack16(*reinterpret_cast<const uint4 *>(s + 16), b);
        unpack16(*reinterpret_cast<const uint4 *>(s + 32), c);
#pragma unroll
        for (int i = 0; i < 16; i++) { px[i] = a[i]; px[16 + i] = b[i]; px[32 + i] = c[i]; }
    } else {
        for (int i = 0; i < 48; i++) px[i] = (i < 3 * nv) ? s[i] : 0u;
    }
    const float wr = 0.2126f, wg = 0.7152f, wb = 0.0722f;
    uint32_t o[16];
#pragma unroll
    for (int i = 0; i < 16; i++) {
        float v = (float)px[3 * i] * wr;
        v = v + (float)px[3 * i + 1] * wg;
        v = v + (float)px[3 * i + 2] * wb;
        int iv = (int)v;
        o[i] = (uint32_t)min(max(iv, 0), 255);
    }
    st16_partial(dst + (long)y * dstStride + x0, o, nv);
}

// ------------------------------------------------------------------ ColorConvert IYUV <- RGB ---
// HafCpu_ColorConvert_IYUV_RGB (ago_haf_cpu_color_convert.cpp:2492-2705).  One thread per 4 columns of a
// row pair.  Columns below w & ~3 follow the SSE body: per pixel, in float (explicit round-to-nearest
// operations, no contraction), Y = ((R*0.2126f) + G*0.7152f) + B*0.0722f, U = (((R*-0.1146f) +
// G*-0.3854f) + B*0.5f) + 128, V = (((R*0.5f) + G*-0.4542f) + B*-0.0458f) + 128, each truncated; each
// chroma sample averages vertically with rounding, then horizontally with rounding.  The last two columns
// when w % 4 == 2 follow the scalar tail: Y adds B*0.0722 in double, and U / V are the float sum of the four
// values (row-major order) divided by 4, truncated.
MVX_DEV float yuv_dot(float r, float g, float b, float wr, float wg, float wb)
{
    return __fadd_rn(__fadd_rn(__fmul_rn(r, wr), __fmul_rn(g, wg)), __fmul_rn(b, wb));
}

__global__ void __launch_bounds__(256)
k_iyuv_from_rgb(uint32_t w, uint32_t h, uint8_t *__restrict__ dy, uint32_t ys, uint8_t *__restrict__ du, uint32_t us,
                uint8_t *__restrict__ dv, uint32_t vs, const uint8_t *__restrict__ rgb, uint32_t rs)
{
    const uint32_t x0 = (blockIdx.x * 256 + threadIdx.x) * 4;
    const uint32_t r = blockIdx.y * 2;
    if (x0 >= w || r >= h) return;
    const bool body = x0 < (w & ~3u);
    const int n = body ? 4 : 2;
    float R[2][4], G[2][4], B[2][4];
#pragma unroll
    for (int q = 0; q < 2; q++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const uint8_t *p = rgb + (long)(r + q) * rs + 3L * (x0 + (i < n ? i : 0));
            R[q][i] = p[0];
            G[q][i] = p[1];
            B[q][i] = p[2];
        }
    if (body) {
        int uu[2][4], vv[2][4];
#pragma unroll
        for (int q = 0; q < 2; q++) {
            uint32_t yw = 0;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int yi = min(max((int)yuv_dot(R[q][i], G[q][i], B[q][i], 0.2126f, 0.7152f, 0.0722f), 0), 255);
                yw |= (uint32_t)yi << (8 * i);
                uu[q][i] = min(max((int)__fadd_rn(yuv_dot(R[q][i], G[q][i], B[q][i], -0.1146f, -0.3854f, 0.5f), 128.0f), 0), 65535);
                vv[q][i] = min(max((int)__fadd_rn(yuv_dot(R[q][i], G[q][i], B[q][i], 0.5f, -0.4542f, -0.0458f), 128.0f), 0), 65535);
            }
            uint8_t *py = dy + (long)(r + q) * ys + x0;
#pragma unroll
            for (int i = 0; i < 4; i++) py[i] = (uint8_t)(yw >> (8 * i));
        }
#pragma unroll
        for (int j = 0; j < 2; j++) {
            const int uo = (((uu[0][2 * j] + uu[1][2 * j] + 1) >> 1) + ((uu[0][2 * j + 1] + uu[1][2 * j + 1] + 1) >> 1) + 1) >> 1;
            const int vo = (((vv[0][2 * j] + vv[1][2 * j] + 1) >> 1) + ((vv[0][2 * j + 1] + vv[1][2 * j + 1] + 1) >> 1) + 1) >> 1;
            du[(long)blockIdx.y * us + (x0 >> 1) + j] = (uint8_t)min(uo, 255);
            dv[(long)blockIdx.y * vs + (x0 >> 1) + j] = (uint8_t)min(vo, 255);
        }
    } else {
        float U = 0.0f, V = 0.0f;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int q = k >> 1, i = k & 1;
            const float rg = __fadd_rn(__fmul_rn(R[q][i], 0.2126f), __fmul_rn(G[q][i], 0.7152f));
            const double yd = __dadd_rn((double)rg, __dmul_rn((double)B[q][i], 0.0722));
            dy[(long)(r + q) * ys + x0 + i] = (uint8_t)(int)yd;
            const float tu = __fadd_rn(yuv_dot(R[q][i], G[q][i], B[q][i], -0.1146f, -0.3854f, 0.5f), 128.0f);
            const float tv = __fadd_rn(yuv_dot(R[q][i], G[q][i], B[q][i], 0.5f, -0.4542f, -0.0458f), 128.0f);
            U = k ? __fadd_rn(U, tu) : tu;
            V = k ? __fadd_rn(V, tv) : tv;
        }
        du[(long)blockIdx.y * us + (x0 >> 1)] = (uint8_t)(int)__fmul_rn(U, 0.25f);
        dv[(long)blockIdx.y * vs + (x0 >> 1)] = (uint8_t)(int)__fmul_rn(V, 0.25f);
    }
}

// ------------------------------------------------------------------ ChannelCopy ----------------
__global__ void __launch_bounds__(256)
k_copy_u8(uint32_t w, uint32_t h, uint8_t *__restrict__ dst, uint32_t dstStride, const uint8_t *__restrict__ src,
          uint32_t srcStride)
{
    const uint32_t x0 = (blockIdx.x * 256 + threadIdx.x) * 16;
    const uint32_t y = blockIdx.y;
    if (x0 >= w || y >= h) return;
    const uint8_t *s = src + (long)y * srcStride + x0;
    uint8_t *d = dst + (long)y * dstStride + x0;
    if (x0 + 16 <= w) {
        *reinterpret_cast<uint4 *>(d) = *reinterpret_cast<const uint4 *>(s);
    } else {
        for (uint32_t i = 0; x0 + i < w; i++) d[i] = s[i];
    }
}

// ------------------------------------------------------------------ ScaleImage bilinear --------
// HafCpu_ScaleImage_U8_U8_Bilinear (ago_haf_cpu_geometric.cpp:1977-2105) with the scale matrix of
// ago_kernel_api.cpp:19897-19916; fixed point with 18 fractional bits, 8-bit weights.
struct ScaleParams { int xinc, yinc, xoffs, yoffs; };

__global__ void __launch_bounds__(256)
k_scale_bilinear(uint32_t dw, uint32_t dh, uint8_t *__restrict__ dst, uint32_t dstStride, uint32_t sw, uint32_t sh,
                 const uint8_t *__restrict__ src, uint32_t srcStride, ScaleParams sp)
{
    const uint32_t x0 = (blockIdx.x * 256 + threadIdx.x) * 8;
    const uint32_t y = blockIdx.y;
    if (x0 >= dw || y >= dh) return;
    const long n = (long)sh * srcStride;
    const int ypos = sp.yoffs + (int)y * sp.yinc;
    const int ym = ypos >> 18;
    const int yf = ((ypos & 0x3ffff) + 0x200) >> 10;
    const int omy = 0x100 - yf;
    long r0, r1;
    if (ym < 0) { r0 = r1 = 0; }
    else if (ym >= (int)sh - 1) { r0 = r1 = (long)(sh - 1) * srcStride; }
    else { r0 = (long)ym * srcStride; r1 = r0 + srcStride; }
    const uint32_t body = dw & ~7u;
    uint32_t o[8];
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const int x = (int)x0 + i;
        const int xpos = sp.xoffs + x * sp.xinc;
        int xm = xpos >> 18;
        if (xm < 0) xm = 0;
        const int xf = ((xpos & 0x3ffff) + 0x200) >> 10;
        const int omx = 0x100 - xf;
        const int p00 = (int)ld_u8(src, n, r0 + xm), p01 = (int)ld_u8(src, n, r0 + xm + 1);
        const int p10 = (int)ld_u8(src, n, r1 + xm), p11 = (int)ld_u8(src, n, r1 + xm + 1);
        int v;
        if ((uint32_t)x < body) {
            const int t0 = (omx * p00 + xf * p01 + 0x80) >> 8;
            const int t1 = (omx * p10 + xf * p11 + 0x80) >> 8;
            v = min((omy * t0 + yf * t1 + 0x80) >> 8, 255);
        } else {
            v = (omx * omy * p00 + xf * omy * p01 + omx * yf * p10 + xf * yf * p11 + 0x8000) >> 16;
        }
        o[i] = (uint32_t)v;
    }
    uint8_t *d = dst + (long)y * dstStride + x0;
    if (x0 + 8 <= dw) {
        uint2 v;
        v.x = o[0] | (o[1] << 8) | (o[2] << 16) | (o[3] << 24);
        v.y = o[4] | (o[5] << 8) | (o[6] << 16) | (o[7] << 24);
        *reinterpret_cast<uint2 *>(d) = v;
    } else {
        for (uint32_t i = 0; x0 + i < dw; i++) d[i] = (uint8_t)o[i];
    }
}

// the scale matrix of ago_kernel_api.cpp:19897-19916 (half = 0.5: offsets ratio/2 - 1/2) or of the
// nearest form :19790-19793 (half = 0: offsets ratio/2), in fixed point with 18 fractional bits
ScaleParams scale_params(uint32_t dw, uint32_t dh, uint32_t sw, uint32_t sh, double half = 0.5)
{
    const float xscale = (float)((double)sw / (double)dw);
    const float yscale = (float)((double)sh / (double)dh);
    const float xoffset = (float)((double)sw / (double)dw * 0.5 - half);
    const float yoffset = (float)((double)sh / (double)dh * 0.5 - half);
    ScaleParams p;
    p.xinc = (int)((float)(1 << 18) * xscale);
    p.yinc = (int)((float)(1 << 18) * yscale);
    p.xoffs = (int)((float)(1 << 18) * xoffset);
    p.yoffs = (int)((float)(1 << 18) * yoffset);
    return p;
}

// ------------------------------------------------------------------ ScaleImage nearest ---------
// HafCpu_ScaleImage_U8_U8_Nearest (ago_haf_cpu_geometric.cpp:1874-1976), offsets ratio/2
// (ago_kernel_api.cpp:19790-19793): map = pos >> 18 clamped to the source, per axis.
__global__ void __launch_bounds__(256)
k_scale_nearest(uint32_t dw, uint32_t dh, uint8_t *__restrict__ dst, uint32_t dstStride, uint32_t sw, uint32_t sh,
                const uint8_t *__restrict__ src, uint32_t srcStride, ScaleParams sp)
{
    const uint32_t x0 = (blockIdx.x * 256 + threadIdx.x) * 8;
    const uint32_t y = blockIdx.y;
    if (x0 >= dw || y >= dh) return;
    const int ym = min(max((int)((uint32_t)sp.yoffs + y * (uint32_t)sp.yinc) >> 18, 0), (int)sh - 1);
    const uint8_t *row = src + (long)ym * srcStride;
    uint32_t o[8];
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const int xm = min(max((int)((uint32_t)sp.xoffs + (x0 + i) * (uint32_t)sp.xinc) >> 18, 0), (int)sw - 1);
        o[i] = (x0 + i < dw) ? row[xm] : 0u;
    }
    uint8_t *d = dst + (long)y * dstStride + x0;
    if (x0 + 8 <= dw) {
        uint2 v;
        v.x = o[0] | (o[1] << 8) | (o[2] << 16) | (o[3] << 24);
        v.y = o[4] | (o[5] << 8) | (o[6] << 16) | (o[7] << 24);
        *reinterpret_cast<uint2 *>(d) = v;
    } else {
        for (uint32_t i = 0; x0 + i < dw; i++) d[i] = (uint8_t)o[i];
    }
}

// ------------------------------------------------------------------ ScaleImage bilinear, borders -
// HafCpu_ScaleImage_U8_U8_Bilinear_Replicate (ago_haf_cpu_geometric.cpp:2108-2258) and _Constant
// (:2260-2391).  Rows: ym < 0 -> (row 0, row 0) | (border, row 0); ym >= sh-1 -> (sh-1, sh-1) |
// (sh-1, border).  Columns below `split` (computed on the host exactly as the reference's
// newDstWidth) take the SIMD two-stage rounding; the others the single rounding with the right
// neighbour at the last source column replaced (p0 | border); on an upscale with split > 0 column 0
// is recomputed with the left neighbour replaced.
template <bool CONSTANT>
__global__ void __launch_bounds__(256)
k_scale_bilinear_border(uint32_t dw, uint32_t dh, uint8_t *__restrict__ dst, uint32_t dstStride, uint32_t sw,
                        uint32_t sh, const uint8_t *__restrict__ src, uint32_t srcStride, ScaleParams sp, int split,
                        int upscale, int border)
{
    const uint32_t x0 = (blockIdx.x * 256 + threadIdx.x) * 8;
    const uint32_t y = blockIdx.y;
    if (x0 >= dw || y >= dh) return;
    const int ypos = (int)((uint32_t)sp.yoffs + y * (uint32_t)sp.yinc);
    const int ym = ypos >> 18;
    const int yf = ((ypos & 0x3ffff) + 0x200) >> 10, omy = 0x100 - yf;
    const uint8_t *r1, *r2;   // nullptr: the border row
    if (ym < 0) { r1 = CONSTANT ? nullptr : src; r2 = src; }
    else if (ym >= (int)sh - 1) { r1 = src + (long)(sh - 1) * srcStride; r2 = CONSTANT ? nullptr : r1; }
    else { r1 = src + (long)ym * srcStride; r2 = r1 + srcStride; }
    auto px = [&](const uint8_t *r, int c) -> int { return (CONSTANT && !r) ? border : (int)r[c]; };
    uint32_t o[8];
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const int x = (int)x0 + i;
        const int xpos = (int)((uint32_t)sp.xoffs + (uint32_t)x * (uint32_t)sp.xinc);
        int xm = xpos >> 18;
        xm = xm >= (int)sw - 1 ? (int)sw - 1 : max(xm, 0);
        const int xf = ((xpos & 0x3ffff) + 0x200) >> 10, omx = 0x100 - xf;
        int v = 0;
        if ((uint32_t)x < dw) {
            if (x == 0 && upscale && split > 0) {
                // column 0 recomputed after the SIMD pass (column 0 of both rows; xm is 0 here)
                int p0, p1, p2, p3;
                if (CONSTANT) { p0 = p2 = border; p1 = px(r1, 0); p3 = px(r2, 0); }
                else { p0 = p1 = px(r1, 0); p2 = p3 = px(r2, 0); }
                v = (int)(((uint32_t)(omx * omy * p0) + (uint32_t)(xf * omy * p1) + (uint32_t)(omx * yf * p2) +
                           (uint32_t)(xf * yf * p3) + 0x8000u) >> 16);
            } else if (x < split) {
                // below the split every map is below sw-1, so xm + 1 stays inside the row
                const int t1 = (omx * px(r1, xm) + xf * px(r1, xm + 1) + 0x80) >> 8;
                const int t2 = (omx * px(r2, xm) + xf * px(r2, xm + 1) + 0x80) >> 8;
                v = (omy * t1 + yf * t2 + 0x80) >> 8;
            } else {
                const int p0 = px(r1, xm), p2 = px(r2, xm);
                const bool last = xm >= (int)sw - 1;
                const int p1 = last ? (CONSTANT ? border : p0) : px(r1, xm + 1);
                const int p3 = last ? (CONSTANT ? border : p2) : px(r2, xm + 1);
                v = (int)(((uint32_t)(omx * omy * p0) + (uint32_t)(xf * omy * p1) + (uint32_t)(omx * yf * p2) +
                           (uint32_t)(xf * yf * p3) + 0x8000u) >> 16);
            }
        }
        o[i] = (uint32_t)v;
    }
    uint8_t *d = dst + (long)y * dstStride + x0;
    if (x0 + 8 <= dw) {
        uint2 v;
        v.x = o[0] | (o[1] << 8) | (o[2] << 16) | (o[3] << 24);
        v.y = o[4] | (o[5] << 8) | (o[6] << 16) | (o[7] << 24);
        *reinterpret_cast<uint2 *>(d) = v;
    } else {
        for (uint32_t i = 0; x0 + i < dw; i++) d[i] = (uint8_t)o[i];
    }
}

// ------------------------------------------------------------------ ScaleImage area (2x) --------
// HafCpu_ScaleImage_U8_U8_Area, exact 2x form (ago_haf_cpu_geometric.cpp:2518-2545, offsets -1/2):
// output row y averages source rows S = (int)(2y - 0.5f) clamped to sh-2, and S+1, in 2x2 blocks,
// (a + b + c + d + 2) >> 2, columns below dw & ~7 (the rest are not written).  One thread: 8
// outputs from 16 source bytes of each row (16-byte aligned: rows are, and 2x is a multiple of 16).
__global__ void __launch_bounds__(256)
k_scale_area2(uint32_t dw8, uint32_t dh, uint8_t *__restrict__ dst, uint32_t dstStride, uint32_t sh,
              const uint8_t *__restrict__ src, uint32_t srcStride)
{
    const uint32_t x0 = (blockIdx.x * 256 + threadIdx.x) * 8;
    const uint32_t y = blockIdx.y;
    if (x0 >= dw8 || y >= dh) return;
    int r = (int)(-0.5f + (float)(2u * y));
    r = min(max(r, 0), (int)sh - 2);
    const uint4 a = ld16(src + (long)r * srcStride + 2 * x0), b = ld16(src + (long)(r + 1) * srcStride + 2 * x0);
    uint32_t w0 = 0, w1 = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const uint32_t v = (byte_of(a, 2 * i) + byte_of(a, 2 * i + 1) + byte_of(b, 2 * i) + byte_of(b, 2 * i + 1) + 2u) >> 2;
        if (i < 4) w0 |= v << (8 * i);
        else w1 |= v << (8 * (i - 4));
    }
    uint2 v;
    v.x = w0; v.y = w1;
    *reinterpret_cast<uint2 *>(dst + (long)y * dstStride + x0) = v;
}

// ------------------------------------------------------------------ ScaleGaussianHalf 5x5 ------
// one thread: 8 outputs x PYR_ROWS rows (pyramid.hpp)
__global__ void __launch_bounds__(256)
k_gaussian_half_5x5(uint32_t dw, uint32_t dh, uint8_t *__restrict__ dst, uint32_t dstStride, uint32_t sw, uint32_t sh,
                    const uint8_t *__restrict__ src, uint32_t srcStride)
{
    const uint32_t x0 = (blockIdx.x * 64 + threadIdx.x) * 8;
    const int y0 = 1 + (int)(blockIdx.y * 4 + threadIdx.y) * PYR_ROWS;
    half_strip(dw, dh, dst, dstStride, sw, sh, src, srcStride, x0, y0);
}

// ------------------------------------------------------------------ ScaleGaussianHalf 3x3 ------
// HafCpu_ScaleGaussianHalf_U8_U8_3x3 (ago_haf_cpu_geometric.cpp:2772-2829) over output rows 1..dh-2
// (ago_kernel_api.cpp:15348): output row y is centred on source row 2y; V(c) = [1 2 1] down the three
// rows, V(-1) = 0 (the reference's zeroed scratch word), out = (V(2x-1) + 2V(2x) + V(2x+1)) >> 4.
// One thread: 8 outputs of one row from one aligned 16-byte load per source row (the host checks
// dw <= (sw+1)/2 and dh <= (sh+1)/2, which keeps the loads inside the row strides).
__global__ void __launch_bounds__(256)
k_gaussian_half_3x3(uint32_t dw, uint32_t dh, uint8_t *__restrict__ dst, uint32_t dstStride,
                    const uint8_t *__restrict__ src, uint32_t srcStride)
{
    const uint32_t x0 = (blockIdx.x * 256 + threadIdx.x) * 8;
    const uint32_t y = blockIdx.y + 1;
    if (x0 >= dw || y + 1 >= dh) return;
    const uint8_t *r0 = src + (long)(2 * y - 1) * srcStride + 2 * x0;
    uint32_t a[16], b[16], c[16];
    unpack16(ld16(r0), a);
    unpack16(ld16(r0 + srcStride), b);
    unpack16(ld16(r0 + 2 * (long)srcStride), c);
    const uint32_t vm1 = x0 ? (uint32_t)r0[-1] + 2u * r0[(long)srcStride - 1] + r0[2 * (long)srcStride - 1] : 0u;
    uint32_t v[16];
#pragma unroll
    for (int i = 0; i < 16; i++) v[i] = a[i] + 2u * b[i] + c[i];
    uint32_t o[16];
#pragma unroll
    for (int k = 0; k < 8; k++) o[k] = ((k ? v[2 * k - 1] : vm1) + 2u * v[2 * k] + v[2 * k + 1]) >> 4;
    uint8_t *d = dst + (long)y * dstStride + x0;
    if (x0 + 8 <= dw) {
        uint2 w;
        w.x = o[0] | (o[1] << 8) | (o[2] << 16) | (o[3] << 24);
        w.y = o[4] | (o[5] << 8) | (o[6] << 16) | (o[7] << 24);
        *reinterpret_cast<uint2 *>(d) = w;
    } else {
        for (uint32_t k = 0; k < dw - x0; k++) d[k] = (uint8_t)o[k];
    }
}

// ------------------------------------------------------------------ ChannelExtract U8 <- U16 / U24 ---
// HafCpu_ChannelExtract_U8_U16_Pos0/1 and _U8_U24_Pos0/1/2 (ago_haf_cpu_ch_extract_combine.cpp:663 on):
// byte `pos` of every `bpp`-byte pixel (NV12 / NV21 chroma samples, RGB channels).  One thread: 16 outputs.
__global__ void __launch_bounds__(256)
k_extract_u8(uint32_t w, uint32_t h, uint8_t *__restrict__ dst, uint32_t dstStride, const uint8_t *__restrict__ src,
             uint32_t srcStride, uint32_t bpp, uint32_t pos)
{
    const uint32_t x0 = (blockIdx.x * 256 + threadIdx.x) * 16, y = blockIdx.y;
    if (x0 >= w || y >= h) return;
    const uint8_t *s = src + (size_t)y * srcStride + (size_t)bpp * x0 + pos;
    uint8_t *d = dst + (size_t)y * dstStride + x0;
    const uint32_t n = min(16u, w - x0);
    for (uint32_t k = 0; k < n; k++) d[k] = s[bpp * k];
}

// ------------------------------------------------------------------ ColorConvert RGB <- YUV ------
// HafCpu_ColorConvert_RGB_IYUV / _RGB_NV12 (ago_haf_cpu_color_convert.cpp:472-768), BT.709: one chroma
// sample per 2x2 block, u = U - 128, v = V - 128; R = Y + v*1.5748, G = Y - (u*0.1873 + v*0.4681),
// B = Y + u*1.8556, each truncated and clamped to [0, 255].  One thread: 4 consecutive pixels of a row.
template <bool NV12>
__global__ void __launch_bounds__(256)
k_rgb_from_yuv(uint32_t w, uint32_t h, uint8_t *__restrict__ dst, uint32_t dstStride, const uint8_t *__restrict__ ys, uint32_t yStride,
               const uint8_t *__restrict__ us, uint32_t uStride, const uint8_t *__restrict__ vs, uint32_t vStride)
{
    const uint32_t x0 = (blockIdx.x * 256 + threadIdx.x) * 4, r = blockIdx.y;
    if (x0 >= w || r >= h) return;
    uint8_t *d = dst + (size_t)r * dstStride + 3ull * x0;
#pragma unroll
    for (uint32_t k = 0; k < 4; k++) {
        const uint32_t x = x0 + k;
        if (x >= w) break;
        const float Y = (float)ys[(size_t)r * yStride + x];
        float U, V;
        if (NV12) {
            const uint8_t *c = us + (size_t)(r / 2) * uStride + 2 * (x / 2);
            U = (float)c[0] - 128.0f;
            V = (float)c[1] - 128.0f;
        } else {
            U = (float)us[(size_t)(r / 2) * uStride + x / 2] - 128.0f;
            V = (float)vs[(size_t)(r / 2) * vStride + x / 2] - 128.0f;
        }
        const float G = (U * 0.1873f) + (V * 0.4681f), R = V * 1.5748f, B = U * 1.8556f;
        d[3 * k + 0] = (uint8_t)fminf(fmaxf(Y + R, 0.0f), 255.0f);
        d[3 * k + 1] = (uint8_t)fminf(fmaxf(Y - G, 0.0f), 255.0f);
        d[3 * k + 2] = (uint8_t)fminf(fmaxf(Y + B, 0.0f), 255.0f);
    }
}

}  // namespace mvx

using namespace mvx;

namespace {
int rgb_from_yuv(mvx_stream_t stream, uint32_t w, uint32_t h, uint8_t *dst, uint32_t dstStride, const uint8_t *y, uint32_t yStride,
                 const uint8_t *u, uint32_t uStride, const uint8_t *v, uint32_t vStride, bool nv12)
{
    if (w == 0 || h == 0) return MVX_SUCCESS;
    if ((w & 1u) || (h & 1u)) return -15;   // the reference converts 2x2 blocks and writes past odd sizes
    if (!dst || !y || !u || (!nv12 && !v) || dstStride < 3 * w || yStride < w || uStride < (nv12 ? w : w / 2) ||
        (!nv12 && vStride < w / 2))
        return MVX_ERROR_INVALID_PARAMETERS;
    dim3 grid(((w + 3) / 4 + 255) / 256, h);
    if (nv12)
        hipLaunchKernelGGL(k_rgb_from_yuv<true>, grid, dim3(256), 0, stream_of(stream), w, h, dst, dstStride, y, yStride, u, uStride, u,
                           uStride);
    else
        hipLaunchKernelGGL(k_rgb_from_yuv<false>, grid, dim3(256), 0, stream_of(stream), w, h, dst, dstStride, y, yStride, u, uStride, v,
                           vStride);
    MVX_LAUNCH_CHECK();
    return MVX_SUCCESS;
}
}  // namespace

extern "C" int mvx_ChannelExtract_U8_U16(mvx_stream_t stream, uint32_t w, uint32_t h, uint8_t *dst, uint32_t dstStride, const uint8_t *src,
                                         uint32_t srcStride, uint32_t pos)
{
    if (w == 0 || h == 0) return MVX_SUCCESS;
    if (!dst || !src || pos > 1 || dstStride < w || srcStride < 2 * w) return MVX_ERROR_INVALID_PARAMETERS;
    dim3 grid(((w + 15) / 16 + 255) / 256, h);
    hipLaunchKernelGGL(k_extract_u8, grid, dim3(256), 0, stream_of(stream), w, h, dst, dstStride, src, srcStride, 2u, pos);
    MVX_LAUNCH_CHECK();
    return MVX_SUCCESS;
}

extern "C" int mvx_ChannelExtract_U8_U24(mvx_stream_t stream, uint32_t w, uint32_t h, uint8_t *dst, uint32_t dstStride, const uint8_t *src,
                                         uint32_t srcStride, uint32_t pos)
{
    if (w == 0 || h == 0) return MVX_SUCCESS;
    if (!dst || !src || pos > 2 || dstStride < w || srcStride < 3 * w) return MVX_ERROR_INVALID_PARAMETERS;
    dim3 grid(((w + 15) / 16 + 255) / 256, h);
    hipLaunchKernelGGL(k_extract_u8, grid, dim3(256), 0, stream_of(stream), w, h, dst, dstStride, src, srcStride, 3u, pos);
    MVX_LAUNCH_CHECK();
    return MVX_SUCCESS;
}

extern "C" int mvx_ColorConvert_RGB_IYUV(mvx_stream_t stream, uint32_t w, uint32_t h, uint8_t *dst, uint32_t dstStride, const uint8_t *y,
                                         uint32_t yStride, const uint8_t *u, uint32_t uStride, const uint8_t *v, uint32_t vStride)
{
    return rgb_from_yuv(stream, w, h, dst, dstStride, y, yStride, u, uStride, v, vStride, false);
}

extern "C" int mvx_ColorConvert_RGB_NV12(mvx_stream_t stream, uint32_t w, uint32_t h, uint8_t *dst, uint32_t dstStride, const uint8_t *y,
                                         uint32_t yStride, const uint8_t *uv, uint32_t uvStride)
{
    return rgb_from_yuv(stream, w, h, dst, dstStride, y, yStride, uv, uvStride, nullptr, 0, true);
}

extern "C" int mvx_ColorConvert_Y_RGB(mvx_stream_t stream, uint32_t w, uint32_t h, uint8_t *dst, uint32_t dstStride,
                                      const uint8_t *rgb, uint32_t rgbStride)
{
    if (w == 0 || h == 0) return MVX_SUCCESS;
    if (!plane_ok_g(dst, w, dstStride, 1) || !plane_ok_g(rgb, w, rgbStride, 3)) return MVX_ERROR_INVALID_PARAMETERS;
    dim3 grid(((w + 15) / 16 + 255) / 256, h);
    hipLaunchKernelGGL(k_color_y_rgb, grid, dim3(256), 0, stream_of(stream), w, h, dst, dstStride, rgb, rgbStride);
    MVX_LAUNCH_CHECK();
    return MVX_SUCCESS;
}

extern "C" int mvx_ColorConvert_IYUV_RGB(mvx_stream_t stream, uint32_t w, uint32_t h, uint8_t *dstY, uint32_t yStride, uint8_t *dstU,
                                         uint32_t uStride, uint8_t *dstV, uint32_t vStride, const uint8_t *rgb, uint32_t rgbStride)
{
    if (w == 0 || h == 0) return MVX_SUCCESS;
    if ((w & 1u) || (h & 1u)) return -15;
    if (!plane_ok_g(dstY, w, yStride, 1) || !plane_ok_g(dstU, w / 2, uStride, 1) || !plane_ok_g(dstV, w / 2, vStride, 1) ||
        !plane_ok_g(rgb, w, rgbStride, 3))
        return MVX_ERROR_INVALID_PARAMETERS;
    dim3 grid(((w + 3) / 4 + 255) / 256, h / 2);
    hipLaunchKernelGGL(k_iyuv_from_rgb, grid, dim3(256), 0, stream_of(stream), w, h, dstY, yStride, dstU, uStride, dstV, vStride, rgb,
                       rgbStride);
    MVX_LAUNCH_CHECK();
    return MVX_SUCCESS;
}

extern "C" int mvx_ChannelCopy_U8_U8(mvx_stream_t stream, uint32_t w, uint32_t h, uint8_t *dst, uint32_t dstStride,
                                     const uint8_t *src, uint32_t srcStride)
{
    if (w == 0 || h == 0) return MVX_SUCCESS;
    if (!plane_ok_g(dst, w, dstStride, 1) || !plane_ok_g(src, w, srcStride, 1)) return MVX_ERROR_INVALID_PARAMETERS;
    dim3 grid(((w + 15) / 16 + 255) / 256, h);
    hipLaunchKernelGGL(k_copy_u8, grid, dim3(256), 0, stream_of(stream), w, h, dst, dstStride, src, srcStride);
    MVX_LAUNCH_CHECK();
    return MVX_SUCCESS;
}

extern "C" int mvx_ScaleImage_U8_U8_Bilinear(mvx_stream_t stream, uint32_t dw, uint32_t dh, uint8_t *dst,
                                             uint32_t dstStride, uint32_t sw, uint32_t sh, const uint8_t *src,
                                             uint32_t srcStride)
{
    if (dw == 0 || dh == 0) return MVX_SUCCESS;
    if (sw == 0 || sh == 0 || !plane_ok_g(dst, dw, dstStride, 1) || !plane_ok_g(src, sw, srcStride, 1) || (dstStride & 7u))
        return MVX_ERROR_INVALID_PARAMETERS;
    ScaleParams sp = scale_params(dw, dh, sw, sh);
    dim3 grid(((dw + 7) / 8 + 255) / 256, dh);
    hipLaunchKernelGGL(k_scale_bilinear, grid, dim3(256), 0, stream_of(stream), dw, dh, dst, dstStride, sw, sh, src,
                       srcStride, sp);
    MVX_LAUNCH_CHECK();
    return MVX_SUCCESS;
}

extern "C" int mvx_ScaleImage_U8_U8_Nearest(mvx_stream_t stream, uint32_t dw, uint32_t dh, uint8_t *dst,
                                           uint32_t dstStride, uint32_t sw, uint32_t sh, const uint8_t *src,
                                           uint32_t srcStride)
{
    if (dw == 0 || dh == 0) return MVX_SUCCESS;
    if (sw == 0 || sh == 0 || sw > 65535 || sh > 65535 || !plane_ok_g(dst, dw, dstStride, 1) ||
        !plane_ok_g(src, sw, srcStride, 1) || (dstStride & 7u))
        return MVX_ERROR_INVALID_PARAMETERS;
    ScaleParams sp = scale_params(dw, dh, sw, sh, 0.0);
    dim3 grid(((dw + 7) / 8 + 255) / 256, dh);
    hipLaunchKernelGGL(k_scale_nearest, grid, dim3(256), 0, stream_of(stream), dw, dh, dst, dstStride, sw, sh, src,
                       srcStride, sp);
    MVX_LAUNCH_CHECK();
    return MVX_SUCCESS;
}

namespace {
// newDstWidth of the bilinear border forms (ago_haf_cpu_geometric.cpp:2133-2155): (first column whose
// map reaches sw-1) - 1, the value 0 counting as unset so a later column sets it again; 0 when none
// does; rounded down to a multiple of 8 when dw is not one
int border_split(uint32_t dw, uint32_t sw, const ScaleParams &sp)
{
    uint32_t nd = 0;
    for (uint32_t x = 0; x < dw; x++)
        if (((int)((uint32_t)sp.xoffs + x * (uint32_t)sp.xinc) >> 18) >= (int)sw - 1 && !nd) nd = x - 1u;
    if (dw & 7u) nd &= ~7u;
    return (int)nd;
}

int scale_bilinear_border(mvx_stream_t stream, uint32_t dw, uint32_t dh, uint8_t *dst, uint32_t dstStride, uint32_t sw,
                          uint32_t sh, const uint8_t *src, uint32_t srcStride, bool constant, uint8_t border)
{
    if (dw == 0 || dh == 0) return MVX_SUCCESS;
    // a source of one column makes the reference's column split wrap below zero (it then runs its
    // 8-wide loop past the row); widths of at least 2 keep every access inside the source row
    if (sw < 2 || sh == 0 || !plane_ok_g(dst, dw, dstStride, 1) || !plane_ok_g(src, sw, srcStride, 1) || (dstStride & 7u))
        return MVX_ERROR_INVALID_PARAMETERS;
    const ScaleParams sp = scale_params(dw, dh, sw, sh);
    const int split = border_split(dw, sw, sp);
    const int upscale = (float)((double)sw / (double)dw) < 1.0f ? 1 : 0;
    dim3 grid(((dw + 7) / 8 + 255) / 256, dh);
    if (constant)
        hipLaunchKernelGGL(k_scale_bilinear_border<true>, grid, dim3(256), 0, stream_of(stream), dw, dh, dst, dstStride,
                           sw, sh, src, srcStride, sp, split, upscale, (int)border);
    else
        hipLaunchKernelGGL(k_scale_bilinear_border<false>, grid, dim3(256), 0, stream_of(stream), dw, dh, dst, dstStride,
                           sw, sh, src, srcStride, sp, split, upscale, 0);
    MVX_LAUNCH_CHECK();
    return MVX_SUCCESS;
}
}  // namespace

extern "C" int mvx_ScaleImage_U8_U8_Bilinear_Replicate(mvx_stream_t stream, uint32_t dw, uint32_t dh, uint8_t *dst,
                                                       uint32_t dstStride, uint32_t sw, uint32_t sh, const uint8_t *src,
                                                       uint32_t srcStride)
{
    return scale_bilinear_border(stream, dw, dh, dst, dstStride, sw, sh, src, srcStride, false, 0);
}

extern "C" int mvx_ScaleImage_U8_U8_Bilinear_Constant(mvx_stream_t stream, uint32_t dw, uint32_t dh, uint8_t *dst,
                                                      uint32_t dstStride, uint32_t sw, uint32_t sh, const uint8_t *src,
                                                      uint32_t srcStride, uint8_t border)
{
    return scale_bilinear_border(stream, dw, dh, dst, dstStride, sw, sh, src, srcStride, true, border);
}

extern "C" int mvx_ScaleImage_U8_U8_Area(mvx_stream_t stream, uint32_t dw, uint32_t dh, uint8_t *dst,
                                         uint32_t dstStride, uint32_t sw, uint32_t sh, const uint8_t *src,
                                         uint32_t srcStride)
{
    if (dw == 0 || dh == 0) return MVX_SUCCESS;
    if (sw == 0 || sh == 0 || !plane_ok_g(dst, dw, dstStride, 1) || !plane_ok_g(src, sw, srcStride, 1) || (dstStride & 7u))
        return MVX_ERROR_INVALID_PARAMETERS;
    // the reference's graph optimizer turns a same-size scale into a plane copy and an area scale that
    // does not shrink both dimensions into nearest neighbour (ago_drama_remove.cpp:1270-1311)
    if (sw == dw && sh == dh) return mvx_ChannelCopy_U8_U8(stream, dw, dh, dst, dstStride, src, srcStride);
    if (!(sw > dw && sh > dh)) return mvx_ScaleImage_U8_U8_Nearest(stream, dw, dh, dst, dstStride, sw, sh, src, srcStride);
    const float xs = (float)((double)sw / (double)dw), ys = (float)((double)sh / (double)dh);
    if (xs == 2.0f && ys == 2.0f && dw >= 8 && sh >= 2) {
        const uint32_t dw8 = dw & ~7u;
        dim3 grid((dw8 / 8 + 255) / 256, dh);
        hipLaunchKernelGGL(k_scale_area2, grid, dim3(256), 0, stream_of(stream), dw8, dh, dst, dstStride, sh, src, srcStride);
        MVX_LAUNCH_CHECK();
        return MVX_SUCCESS;
    }
    // the reference's generic area path keeps its row sums past the end of its node buffer, so its
    // result is not defined (ago_kernel_api.cpp:20269 vs ago_haf_cpu_geometric.cpp:2553)
    return MVX_ERROR_NOT_SUPPORTED;
}

extern "C" int mvx_ScaleGaussianHalf_U8_U8_3x3(mvx_stream_t stream, uint32_t dw, uint32_t dh, uint8_t *dst,
                                               uint32_t dstStride, uint32_t sw, uint32_t sh, const uint8_t *src,
                                               uint32_t srcStride)
{
    // an output row count below 2 wraps the reference's dstHeight - 2; a larger output than half the
    // source (rounded up) would read past the source rows
    if (dh == 1 || dw > (sw + 1) / 2 || dh > (sh + 1) / 2) return -15;
    if (dw == 0 || dh < 3) return MVX_SUCCESS;
    if (!plane_ok_g(dst, dw, dstStride, 1) || !plane_ok_g(src, sw, srcStride, 1) || (dstStride & 7u))
        return MVX_ERROR_INVALID_PARAMETERS;
    dim3 grid(((dw + 7) / 8 + 255) / 256, dh - 2);
    hipLaunchKernelGGL(k_gaussian_half_3x3, grid, dim3(256), 0, stream_of(stream), dw, dh, dst, dstStride, src, srcStride);
    MVX_LAUNCH_CHECK();
    return MVX_SUCCESS;
}

extern "C" int mvx_ScaleGaussianHalf_U8_U8_5x5(mvx_stream_t stream, uint32_t dw, uint32_t dh, uint8_t *dst,
                                               uint32_t dstStride, uint32_t sw, uint32_t sh, const uint8_t *src,
                                               uint32_t srcStride)
{
    // the reference wrapper rejects these sizes with VX_ERROR_INVALID_DIMENSION (ago_kernel_api.cpp:15428-15430)
    if (sw < 5 || sh < 5 || dw < 3 || dh < 3) return -15;
    if (!plane_ok_g(dst, dw, dstStride, 1) || !plane_ok_g(src, sw, srcStride, 1) || (dstStride & 7u))
        return MVX_ERROR_INVALID_PARAMETERS;
    const uint32_t chunks = (dw + 7) / 8;
    const uint32_t strips = (dh - 2 + PYR_ROWS - 1) / PYR_ROWS;
    dim3 grid((chunks + 63) / 64, (strips + 3) / 4);
    hipLaunchKernelGGL(k_gaussian_half_5x5, grid, dim3(64, 4), 0, stream_of(stream), dw, dh, dst, dstStride, sw, sh,
                       src, srcStride);
    MVX_LAUNCH_CHECK();
    return MVX_SUCCESS;
}
