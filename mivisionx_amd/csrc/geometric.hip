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

ScaleParams scale_params(uint32_t dw, uint32_t dh, uint32_t sw, uint32_t sh)
{
    const float xscale = (float)((double)sw / (double)dw);
    const float yscale = (float)((double)sh / (double)dh);
    const float xoffset = (float)((double)sw / (double)dw * 0.5 - 0.5);
    const float yoffset = (float)((double)sh / (double)dh * 0.5 - 0.5);
    ScaleParams p;
    p.xinc = (int)((float)(1 << 18) * xscale);
    p.yinc = (int)((float)(1 << 18) * yscale);
    p.xoffs = (int)((float)(1 << 18) * xoffset);
    p.yoffs = (int)((float)(1 << 18) * yoffset);
    return p;
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

}  // namespace mvx

using namespace mvx;

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
