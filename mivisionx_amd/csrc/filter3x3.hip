// filter3x3.hip -- Gaussian/Box/Median 3x3 (U8 -> U8) and Sobel 3x3 (U8 -> S16 x2) for gfx950.
//
// Work decomposition: one thread owns a 16-pixel column chunk (one 128-bit row load) over a strip
// of ROWS output rows and keeps the three source rows it needs in registers, so every source row
// is fetched from HBM once per strip (plus two halo rows per strip, which stay in L2).  The two
// halo bytes at x0-1 and x0+16 of each row are read at their linear addresses with the zero guard
// of the reference layout, which reproduces the reference CPU's wrap-around at columns 0 and w-1.
//
// Rows 1..h-2 are produced, matching the reference wrappers (ago_kernel_api.cpp:15281-15282 and
// the Sobel/Box/Median equivalents).
#include "mvx_common.hpp"

namespace mvx {

constexpr int F3_BX = 16;     // chunks per block row (16 x 16 px = 256 px)
constexpr int F3_BY = 16;     // strips per block
constexpr int F3_ROWS = 8;    // output rows per thread

struct Row18 { uint32_t p[18]; };   // p[0] = x0-1, p[1..16] = x0..x0+15, p[17] = x0+16

MVX_DEV void load_row18(const uint8_t *src, long n, long base, Row18 &r)
{
    uint32_t v[16];
    unpack16(ld16(src + base), v);
#pragma unroll
    for (int i = 0; i < 16; i++) r.p[i + 1] = v[i];
    r.p[0] = ld_u8(src, n, base - 1);
    r.p[17] = ld_u8(src, n, base + 16);
}

struct OpGaussian {
    MVX_DEV static uint32_t apply(const Row18 &a, const Row18 &b, const Row18 &c, int i)
    {
        uint32_t ha = a.p[i] + 2u * a.p[i + 1] + a.p[i + 2];
        uint32_t hb = b.p[i] + 2u * b.p[i + 1] + b.p[i + 2];
        uint32_t hc = c.p[i] + 2u * c.p[i + 1] + c.p[i + 2];
        return (ha + 2u * hb + hc) >> 4;
    }
};

struct OpBox {
    MVX_DEV static uint32_t apply(const Row18 &a, const Row18 &b, const Row18 &c, int i)
    {
        uint32_t s = a.p[i] + a.p[i + 1] + a.p[i + 2] + b.p[i] + b.p[i + 1] + b.p[i + 2] + c.p[i] + c.p[i + 1] + c.p[i + 2];
        // floor(s / 9) for s <= 2295 as a multiply-high: (s * 7282) >> 16 is exact in that range
        return (s * 7282u) >> 16;
    }
};

MVX_DEV void cswap(uint32_t &a, uint32_t &b) { uint32_t lo = min(a, b), hi = max(a, b); a = lo; b = hi; }

struct OpMedian {
    // exact median of 9 with the 19-exchange selection network of Paeth/Devillard
    MVX_DEV static uint32_t apply(const Row18 &a, const Row18 &b, const Row18 &c, int i)
    {
        uint32_t p0 = a.p[i], p1 = a.p[i + 1], p2 = a.p[i + 2];
        uint32_t p3 = b.p[i], p4 = b.p[i + 1], p5 = b.p[i + 2];
        uint32_t p6 = c.p[i], p7 = c.p[i + 1], p8 = c.p[i + 2];
        cswap(p1, p2); cswap(p4, p5); cswap(p7, p8);
        cswap(p0, p1); cswap(p3, p4); cswap(p6, p7);
        cswap(p1, p2); cswap(p4, p5); cswap(p7, p8);
        cswap(p0, p3); cswap(p5, p8); cswap(p4, p7);
        cswap(p3, p6); cswap(p1, p4); cswap(p2, p5);
        cswap(p4, p7); cswap(p4, p2); cswap(p6, p4);
        cswap(p4, p2);
        return p4;
    }
};

template <class Op>
__global__ void __launch_bounds__(F3_BX * F3_BY)
k_filter3x3_u8(uint32_t w, uint32_t h, uint8_t *__restrict__ dst, uint32_t dstStride,
               const uint8_t *__restrict__ src, uint32_t srcStride)
{
    const uint32_t cx = blockIdx.x * F3_BX + threadIdx.x;
    const uint32_t x0 = cx * 16;
    if (x0 >= w) return;
    const int y0 = 1 + (int)(blockIdx.y * F3_BY + threadIdx.y) * F3_ROWS;
    const int y1 = min(y0 + F3_ROWS, (int)h - 1);
    if (y0 >= y1) return;
    const long n = (long)h * srcStride;
    const int nv = min(16, (int)(w - x0));

    Row18 ra, rb, rc;
    load_row18(src, n, (long)(y0 - 1) * srcStride + x0, ra);
    load_row18(src, n, (long)y0 * srcStride + x0, rb);
    for (int y = y0; y < y1; y++) {
        load_row18(src, n, (long)(y + 1) * srcStride + x0, rc);
        uint32_t o[16];
#pragma unroll
        for (int i = 0; i < 16; i++) o[i] = Op::apply(ra, rb, rc, i);
        st16_partial(dst + (long)y * dstStride + x0, o, nv);
        ra = rb;
        rb = rc;
    }
}

MVX_DEV uint32_t pack2s16(int32_t lo, int32_t hi) { return ((uint32_t)lo & 0xffffu) | ((uint32_t)hi << 16); }

MVX_DEV void st16s16_partial(int16_t *d, const int32_t (&o)[16], int nv)
{
    if (nv == 16) {
        reinterpret_cast<uint4 *>(d)[0] = make_uint4(pack2s16(o[0], o[1]), pack2s16(o[2], o[3]), pack2s16(o[4], o[5]), pack2s16(o[6], o[7]));
        reinterpret_cast<uint4 *>(d)[1] = make_uint4(pack2s16(o[8], o[9]), pack2s16(o[10], o[11]), pack2s16(o[12], o[13]), pack2s16(o[14], o[15]));
    } else {
        for (int i = 0; i < nv; i++) d[i] = (int16_t)o[i];
    }
}

__global__ void __launch_bounds__(F3_BX * F3_BY)
k_sobel3x3(uint32_t w, uint32_t h, int16_t *__restrict__ gx, uint32_t gxStride, int16_t *__restrict__ gy,
           uint32_t gyStride, const uint8_t *__restrict__ src, uint32_t srcStride)
{
    const uint32_t cx = blockIdx.x * F3_BX + threadIdx.x;
    const uint32_t x0 = cx * 16;
    if (x0 >= w) return;
    const int y0 = 1 + (int)(blockIdx.y * F3_BY + threadIdx.y) * F3_ROWS;
    const int y1 = min(y0 + F3_ROWS, (int)h - 1);
    if (y0 >= y1) return;
    const long n = (long)h * srcStride;
    const int nv = min(16, (int)(w - x0));

    Row18 ra, rb, rc;
    load_row18(src, n, (long)(y0 - 1) * srcStride + x0, ra);
    load_row18(src, n, (long)y0 * srcStride + x0, rb);
    for (int y = y0; y < y1; y++) {
        load_row18(src, n, (long)(y + 1) * srcStride + x0, rc);
        int32_t ox[16], oy[16];
#pragma unroll
        for (int i = 0; i < 16; i++) {
            int dxa = (int)ra.p[i + 2] - (int)ra.p[i];
            int dxb = (int)rb.p[i + 2] - (int)rb.p[i];
            int dxc = (int)rc.p[i + 2] - (int)rc.p[i];
            int ha = (int)(ra.p[i] + 2u * ra.p[i + 1] + ra.p[i + 2]);
            int hc = (int)(rc.p[i] + 2u * rc.p[i + 1] + rc.p[i + 2]);
            ox[i] = dxa + 2 * dxb + dxc;
            oy[i] = hc - ha;
        }
        if (gx) st16s16_partial((int16_t *)((uint8_t *)gx + (long)y * gxStride) + x0, ox, nv);
        if (gy) st16s16_partial((int16_t *)((uint8_t *)gy + (long)y * gyStride) + x0, oy, nv);
        ra = rb;
        rb = rc;
    }
}

// stride must hold the row (ALIGN16 of the byte width) and keep 16-byte row alignment
inline bool plane_ok(const void *p, uint32_t w, uint32_t stride, uint32_t bpp)
{
    return p && aligned16(p) && (stride % 16u) == 0 && stride >= ((w * bpp + 15u) & ~15u);
}

template <class Op>
int launch_filter3x3(mvx_stream_t stream, uint32_t w, uint32_t h, uint8_t *dst, uint32_t dstStride,
                     const uint8_t *src, uint32_t srcStride)
{
    if (w == 0 || h == 0) return MVX_SUCCESS;
    if (!plane_ok(dst, w, dstStride, 1) || !plane_ok(src, w, srcStride, 1)) return MVX_ERROR_INVALID_PARAMETERS;
    if (h < 3) return MVX_SUCCESS;
    uint32_t chunks = (w + 15) / 16;
    uint32_t strips = (h - 2 + F3_ROWS - 1) / F3_ROWS;
    dim3 grid((chunks + F3_BX - 1) / F3_BX, (strips + F3_BY - 1) / F3_BY);
    hipLaunchKernelGGL(k_filter3x3_u8<Op>, grid, dim3(F3_BX, F3_BY), 0, stream_of(stream), w, h, dst, dstStride, src, srcStride);
    MVX_LAUNCH_CHECK();
    return MVX_SUCCESS;
}

}  // namespace mvx

using namespace mvx;

extern "C" int mvx_Gaussian_U8_U8_3x3(mvx_stream_t stream, uint32_t w, uint32_t h, uint8_t *dst, uint32_t dstStride,
                                      const uint8_t *src, uint32_t srcStride)
{
    return launch_filter3x3<OpGaussian>(stream, w, h, dst, dstStride, src, srcStride);
}

extern "C" int mvx_Box_U8_U8_3x3(mvx_stream_t stream, uint32_t w, uint32_t h, uint8_t *dst, uint32_t dstStride,
                                 const uint8_t *src, uint32_t srcStride)
{
    return launch_filter3x3<OpBox>(stream, w, h, dst, dstStride, src, srcStride);
}

extern "C" int mvx_Median_U8_U8_3x3(mvx_stream_t stream, uint32_t w, uint32_t h, uint8_t *dst, uint32_t dstStride,
                                    const uint8_t *src, uint32_t srcStride)
{
    return launch_filter3x3<OpMedian>(stream, w, h, dst, dstStride, src, srcStride);
}

extern "C" int mvx_Sobel_S16S16_U8_3x3_GXY(mvx_stream_t stream, uint32_t w, uint32_t h, int16_t *gx, uint32_t gxStride,
                                           int16_t *gy, uint32_t gyStride, const uint8_t *src, uint32_t srcStride)
{
    if (w == 0 || h == 0) return MVX_SUCCESS;
    if (!gx && !gy) return MVX_ERROR_INVALID_PARAMETERS;
    if ((gx && !plane_ok(gx, w, gxStride, 2)) || (gy && !plane_ok(gy, w, gyStride, 2)) || !plane_ok(src, w, srcStride, 1))
        return MVX_ERROR_INVALID_PARAMETERS;
    if (h < 3) return MVX_SUCCESS;
    uint32_t chunks = (w + 15) / 16;
    uint32_t strips = (h - 2 + F3_ROWS - 1) / F3_ROWS;
    dim3 grid((chunks + F3_BX - 1) / F3_BX, (strips + F3_BY - 1) / F3_BY);
    hipLaunchKernelGGL(k_sobel3x3, grid, dim3(F3_BX, F3_BY), 0, stream_of(stream), w, h, gx, gxStride, gy, gyStride, src, srcStride);
    MVX_LAUNCH_CHECK();
    return MVX_SUCCESS;
}
