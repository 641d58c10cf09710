// pyramid.hpp -- one HALF Gaussian pyramid step (ScaleGaussianHalf 5x5) as a device routine, shared by
// the per-node kernel and the frame-batched graph.
//
// HafCpu_ScaleGaussianHalf_U8_U8_5x5 (ago_haf_cpu_pyramid.cpp:94-305), wrapper ago_kernel_api.cpp:15420-15436:
// out(x, y) = (sum_j k_j H(2y+oy-2+j)) >> 8, H(r) = sum_i k_i p(2x+ox-2+i, r), k = [1 4 6 4 1],
// oy = (srcH even), ox = (srcW even), rows 1..dh-2, linear-address taps with the zero guard.
#pragma once
#include "mvx_common.hpp"

namespace mvx {

constexpr int PYR_ROWS = 4;   // output rows per thread

// 16 bytes of the chunk that starts at the 16-aligned byte offset `off`, or zeros outside the plane
MVX_DEV uint4 ld16_guard(const uint8_t *p, long n, long off)
{
    return (off < 0 || off >= n) ? make_uint4(0, 0, 0, 0) : *reinterpret_cast<const uint4 *>(p + off);
}

// horizontal 5-tap at the 8 sample positions 2*x0 + ox + 2i of the source row at byte rowOff
MVX_DEV void pyr_hrow(const uint8_t *src, long n, long rowOff, uint32_t x0, int ox, uint32_t (&H)[8])
{
    const long c = rowOff + 2L * x0;   // 16-aligned
    uint32_t a[16], b[16], m[16];
    unpack16(ld16_guard(src, n, c - 16), m);
    unpack16(ld16_guard(src, n, c), a);
    unpack16(ld16_guard(src, n, c + 16), b);
    // window byte j (j = 0..19) is the source byte at c - 2 + j
    uint32_t win[20];
    win[0] = m[14];
    win[1] = m[15];
#pragma unroll
    for (int j = 0; j < 16; j++) win[2 + j] = a[j];
    win[18] = b[0];
    win[19] = b[1];
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const int k0 = ox + 2 * i;   // centre at window index k0 + 2
        H[i] = win[k0] + 4u * win[k0 + 1] + 6u * win[k0 + 2] + 4u * win[k0 + 3] + win[k0 + 4];
    }
}

// outputs x0..x0+7 of rows y0 .. y0+PYR_ROWS-1 (clipped to 1..dh-2)
MVX_DEV void half_strip(uint32_t dw, uint32_t dh, uint8_t *dst, uint32_t dstStride, uint32_t sw, uint32_t sh,
                        const uint8_t *src, uint32_t srcStride, uint32_t x0, int y0)
{
    if (x0 >= dw) return;
    const int y1 = min(y0 + PYR_ROWS, (int)dh - 1);
    if (y0 >= y1) return;
    const long n = (long)sh * srcStride;
    const int oy = (sh & 1) ? 0 : 1, ox = (sw & 1) ? 0 : 1;
    uint32_t h0[8], h1[8], h2[8], h3[8], h4[8];
    int cy = 2 * y0 + oy;
    pyr_hrow(src, n, (long)(cy - 2) * srcStride, x0, ox, h0);
    pyr_hrow(src, n, (long)(cy - 1) * srcStride, x0, ox, h1);
    pyr_hrow(src, n, (long)cy * srcStride, x0, ox, h2);
    for (int y = y0; y < y1; y++) {
        cy = 2 * y + oy;
        pyr_hrow(src, n, (long)(cy + 1) * srcStride, x0, ox, h3);
        pyr_hrow(src, n, (long)(cy + 2) * srcStride, x0, ox, h4);
        uint32_t o[8];
#pragma unroll
        for (int i = 0; i < 8; i++) o[i] = ((h0[i] + 4u * h1[i] + 6u * h2[i] + 4u * h3[i] + h4[i]) & 0xffffu) >> 8;
        uint8_t *d = dst + (long)y * dstStride + x0;
        if (x0 + 8 <= dw) {
            uint2 v;
            v.x = o[0] | (o[1] << 8) | (o[2] << 16) | (o[3] << 24);
            v.y = o[4] | (o[5] << 8) | (o[6] << 16) | (o[7] << 24);
            *reinterpret_cast<uint2 *>(d) = v;
        } else {
            for (uint32_t i = 0; x0 + i < dw; i++) d[i] = (uint8_t)o[i];
        }
#pragma unroll
        for (int i = 0; i < 8; i++) { h0[i] = h2[i]; h1[i] = h3[i]; h2[i] = h4[i]; }
    }
}

}  // namespace mvx
