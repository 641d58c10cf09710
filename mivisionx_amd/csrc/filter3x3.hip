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
#include "arith.hpp"

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

// Merged Sobel -> Magnitude (+ Phase): the kernels the reference's graph optimizer substitutes when
// Sobel's gx and gy are virtual and read only by Magnitude / Phase (ago_drama_remove.cpp:155-182):
// HafCpu_SobelMagnitudePhase_S16U8_U8_3x3 (ago_haf_cpu_filter.cpp:3972-4202) and
// HafCpu_SobelMagnitude_S16_U8_3x3 (:2468-2682), rows 1..h-2.  Their 16-pixel SSE body covers columns
// [0, w & ~15) and rounds the magnitude half to even (the Magnitude node's body rule); the scalar
// postfix (the last w mod 16 columns) truncates.  A 16-pixel chunk starts at a multiple of 16, so it lies
// wholly in the body or wholly in the postfix.  The phase is the Phase node's per-pixel rule.
template <bool PHASE>
__global__ void __launch_bounds__(F3_BX * F3_BY)
k_sobel_magphase(uint32_t w, uint32_t h, int16_t *__restrict__ mag, uint32_t magStride, uint8_t *__restrict__ phase,
                 uint32_t phaseStride, const uint8_t *__restrict__ src, uint32_t srcStride)
{
    const uint32_t cx = blockIdx.x * F3_BX + threadIdx.x;
    const uint32_t x0 = cx * 16;
    if (x0 >= w) return;
    const int y0 = 1 + (int)(blockIdx.y * F3_BY + threadIdx.y) * F3_ROWS;
    const int y1 = min(y0 + F3_ROWS, (int)h - 1);
    if (y0 >= y1) return;
    const long n = (long)h * srcStride;
    const int nv = min(16, (int)(w - x0));
    const bool tail = x0 >= (w & ~15u);

    Row18 ra, rb, rc;
    load_row18(src, n, (long)(y0 - 1) * srcStride + x0, ra);
    load_row18(src, n, (long)y0 * srcStride + x0, rb);
    for (int y = y0; y < y1; y++) {
        load_row18(src, n, (long)(y + 1) * srcStride + x0, rc);
        int32_t om[16];
        uint32_t op[16];
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const int dxa = (int)ra.p[i + 2] - (int)ra.p[i];
            const int dxb = (int)rb.p[i + 2] - (int)rb.p[i];
            const int dxc = (int)rc.p[i + 2] - (int)rc.p[i];
            const int ha = (int)(ra.p[i] + 2u * ra.p[i + 1] + ra.p[i + 2]);
            const int hc = (int)(rc.p[i] + 2u * rc.p[i + 1] + rc.p[i + 2]);
            const int gx = dxa + 2 * dxb + dxc, gy = hc - ha;
            om[i] = tail ? mag_trunc(gx, gy) : mag_body(gx, gy);
            if (PHASE) op[i] = phase_px(gx, gy);
        }
        st16s16_partial((int16_t *)((uint8_t *)mag + (long)y * magStride) + x0, om, nv);
        if (PHASE) st16_partial(phase + (long)y * phaseStride + x0, op, nv);
        ra = rb;
        rb = rc;
    }
}

// Sobel gradient at (x, y) from linear-address 3x3 reads (the zero guard outside [0, n))
MVX_DEV void grad_at(const uint8_t *src, long n, long stride, int x, int y, int &gx, int &gy)
{
    const long bm = (long)(y - 1) * stride + x, b0 = (long)y * stride + x, bp = (long)(y + 1) * stride + x;
    const int a0 = (int)ld_u8(src, n, bm - 1), a1 = (int)ld_u8(src, n, bm), a2 = (int)ld_u8(src, n, bm + 1);
    const int c0 = (int)ld_u8(src, n, bp - 1), c1 = (int)ld_u8(src, n, bp), c2 = (int)ld_u8(src, n, bp + 1);
    gx = (a2 - a0) + 2 * ((int)ld_u8(src, n, b0 + 1) - (int)ld_u8(src, n, b0 - 1)) + (c2 - c0);
    gy = (c0 + 2 * c1 + c2) - (a0 + 2 * a1 + a2);
}

// HafCpu_SobelPhase_U8_U8_3x3 (ago_haf_cpu_filter.cpp:3946-3970): Sobel GXY into the node's scratch and
// Phase from it, with scratch rows ALIGN16(w) BYTES apart although a row holds w int16 values.  Rows
// are written top to bottom, so each overwrites the second half of the previous one, and Phase reads
// them back at the same spacing: output pixel (x, y) takes the gradient of (x - ALIGN16(w)/2, y + 1)
// when x >= ALIGN16(w)/2 and y + 1 <= h - 2, and its own gradient otherwise.  No scratch is used here:
// each pixel evaluates the gradient the last writer of its scratch cell computed.  One thread per 16
// pixels of a row.
__global__ void __launch_bounds__(256)
k_sobel_phase_merged(uint32_t w, uint32_t h, uint8_t *__restrict__ phase, uint32_t phaseStride, const uint8_t *__restrict__ src,
                     uint32_t srcStride)
{
    const uint32_t x0 = (blockIdx.x * 16 + threadIdx.x) * 16;
    const int y = 1 + (int)(blockIdx.y * 16 + threadIdx.y);
    if (x0 >= w || y > (int)h - 2) return;
    const long n = (long)h * srcStride;
    const int half = (int)(((w + 15u) & ~15u) >> 1);
    const int nv = min(16, (int)(w - x0));
    uint32_t o[16];
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const int x = (int)x0 + i;
        int sx = x, sy = y;
        if (x >= half && y + 1 <= (int)h - 2) { sx = x - half; sy = y + 1; }
        int gx = 0, gy = 0;
        if (i < nv) grad_at(src, n, srcStride, sx, sy, gx, gy);
        o[i] = phase_px(gx, gy);
    }
    st16_partial(phase + (long)y * phaseStride + x0, o, nv);
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

extern "C" int mvx_SobelMagnitudePhase_S16U8_U8_3x3(mvx_stream_t stream, uint32_t w, uint32_t h, int16_t *mag, uint32_t magStride,
                                                    uint8_t *phase, uint32_t phaseStride, const uint8_t *src, uint32_t srcStride)
{
    if (w == 0 || h == 0) return MVX_SUCCESS;
    if (!plane_ok(mag, w, magStride, 2) || (phase && !plane_ok(phase, w, phaseStride, 1)) || !plane_ok(src, w, srcStride, 1))
        return MVX_ERROR_INVALID_PARAMETERS;
    if (h < 3) return MVX_SUCCESS;
    uint32_t chunks = (w + 15) / 16;
    uint32_t strips = (h - 2 + F3_ROWS - 1) / F3_ROWS;
    dim3 grid((chunks + F3_BX - 1) / F3_BX, (strips + F3_BY - 1) / F3_BY);
    if (phase)
        hipLaunchKernelGGL(k_sobel_magphase<true>, grid, dim3(F3_BX, F3_BY), 0, stream_of(stream), w, h, mag, magStride, phase,
                           phaseStride, src, srcStride);
    else
        hipLaunchKernelGGL(k_sobel_magphase<false>, grid, dim3(F3_BX, F3_BY), 0, stream_of(stream), w, h, mag, magStride,
                           (uint8_t *)nullptr, 0u, src, srcStride);
    MVX_LAUNCH_CHECK();
    return MVX_SUCCESS;
}

extern "C" int mvx_SobelMagnitude_S16_U8_3x3(mvx_stream_t stream, uint32_t w, uint32_t h, int16_t *mag, uint32_t magStride,
                                             const uint8_t *src, uint32_t srcStride)
{
    if (w == 0 || h == 0) return MVX_SUCCESS;
    if (!mag) return MVX_ERROR_INVALID_PARAMETERS;
    return mvx_SobelMagnitudePhase_S16U8_U8_3x3(stream, w, h, mag, magStride, nullptr, 0, src, srcStride);
}

extern "C" int mvx_SobelPhase_U8_U8_3x3(mvx_stream_t stream, uint32_t w, uint32_t h, uint8_t *phase, uint32_t phaseStride,
                                        const uint8_t *src, uint32_t srcStride)
{
    if (w == 0 || h == 0) return MVX_SUCCESS;
    if (!plane_ok(phase, w, phaseStride, 1) || !plane_ok(src, w, srcStride, 1)) return MVX_ERROR_INVALID_PARAMETERS;
    if (h < 3) return MVX_SUCCESS;
    const uint32_t chunks = (w + 15) / 16;
    dim3 grid((chunks + 15) / 16, (h - 2 + 15) / 16);
    hipLaunchKernelGGL(k_sobel_phase_merged, grid, dim3(16, 16), 0, stream_of(stream), w, h, phase, phaseStride, src, srcStride);
    MVX_LAUNCH_CHECK();
    return MVX_SUCCESS;
}
