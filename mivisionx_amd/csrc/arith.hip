// arith.hip -- Magnitude (S16,S16 -> S16) and Phase (S16,S16 -> U8) for gfx950.
//
// One thread owns 8 pixels of a row: one 128-bit load of gx and of gy, one 128-bit (Magnitude) or
// 64-bit (Phase) store.  Both are elementwise and HBM-bound (6 B/px and 5 B/px).
// Built with -ffp-contract=off and correctly rounded fp32 divide/sqrt so that float sequences
// follow the reference CPU code operation for operation.
#include "arith.hpp"

namespace mvx {

constexpr int AR_BLOCK = 256;

__global__ void __launch_bounds__(AR_BLOCK)
k_magnitude(uint32_t w, uint32_t h, int16_t *__restrict__ dst, uint32_t dstStride, const int16_t *__restrict__ gx,
            uint32_t gxStride, const int16_t *__restrict__ gy, uint32_t gyStride)
{
    const uint32_t chunks = (w + 7) / 8;
    const uint32_t tid = blockIdx.x * AR_BLOCK + threadIdx.x;
    const uint32_t y = blockIdx.y;
    if (tid >= chunks || y >= h) return;
    const uint32_t x0 = tid * 8;
    const uint32_t body = w & ~7u;
    const int16_t *px = (const int16_t *)((const uint8_t *)gx + (long)y * gxStride) + x0;
    const int16_t *py = (const int16_t *)((const uint8_t *)gy + (long)y * gyStride) + x0;
    int16_t *pd = (int16_t *)((uint8_t *)dst + (long)y * dstStride) + x0;
    // the chunk is inside the 16-byte aligned row (stride >= ALIGN16(2w)), padding lanes read 0
    uint4 vx = *reinterpret_cast<const uint4 *>(px);
    uint4 vy = *reinterpret_cast<const uint4 *>(py);
    const uint32_t ax[4] = {vx.x, vx.y, vx.z, vx.w}, ay[4] = {vy.x, vy.y, vy.z, vy.w};
    int32_t o[8];
    int32_t a[8], b[8];
#pragma unroll
    for (int i = 0; i < 8; i++) {
        a[i] = (int16_t)(ax[i >> 1] >> (16 * (i & 1)));
        b[i] = (int16_t)(ay[i >> 1] >> (16 * (i & 1)));
    }
    if (x0 + 8 <= body) {
#pragma unroll
        for (int i = 0; i < 8; i++) o[i] = mag_body(a[i], b[i]);
    } else {
#pragma unroll
        for (int i = 0; i < 8; i++) o[i] = (x0 + i < body) ? mag_body(a[i], b[i]) : mag_tail(a[i], b[i]);
    }
    if (x0 + 8 <= w) {
        uint4 v;
        v.x = ((uint32_t)o[0] & 0xffff) | ((uint32_t)o[1] << 16);
        v.y = ((uint32_t)o[2] & 0xffff) | ((uint32_t)o[3] << 16);
        v.z = ((uint32_t)o[4] & 0xffff) | ((uint32_t)o[5] << 16);
        v.w = ((uint32_t)o[6] & 0xffff) | ((uint32_t)o[7] << 16);
        *reinterpret_cast<uint4 *>(pd) = v;
    } else {
        for (uint32_t i = 0; x0 + i < w; i++) pd[i] = (int16_t)o[i];
    }
}

__global__ void __launch_bounds__(AR_BLOCK)
k_phase(uint32_t w, uint32_t h, uint8_t *__restrict__ dst, uint32_t dstStride, const int16_t *__restrict__ gx,
        uint32_t gxStride, const int16_t *__restrict__ gy, uint32_t gyStride)
{
    const uint32_t chunks = (w + 7) / 8;
    const uint32_t tid = blockIdx.x * AR_BLOCK + threadIdx.x;
    const uint32_t y = blockIdx.y;
    if (tid >= chunks || y >= h) return;
    const uint32_t x0 = tid * 8;
    const int16_t *px = (const int16_t *)((const uint8_t *)gx + (long)y * gxStride) + x0;
    const int16_t *py = (const int16_t *)((const uint8_t *)gy + (long)y * gyStride) + x0;
    uint8_t *pd = dst + (long)y * dstStride + x0;
    uint4 vx = *reinterpret_cast<const uint4 *>(px);
    uint4 vy = *reinterpret_cast<const uint4 *>(py);
    const uint32_t ax[4] = {vx.x, vx.y, vx.z, vx.w}, ay[4] = {vy.x, vy.y, vy.z, vy.w};
    uint32_t o[8];
#pragma unroll
    for (int i = 0; i < 8; i++) {
        int32_t a = (int16_t)(ax[i >> 1] >> (16 * (i & 1)));
        int32_t b = (int16_t)(ay[i >> 1] >> (16 * (i & 1)));
        o[i] = phase_px(a, b);
    }
    if (x0 + 8 <= w) {
        uint2 v;
        v.x = o[0] | (o[1] << 8) | (o[2] << 16) | (o[3] << 24);
        v.y = o[4] | (o[5] << 8) | (o[6] << 16) | (o[7] << 24);
        *reinterpret_cast<uint2 *>(pd) = v;
    } else {
        for (uint32_t i = 0; x0 + i < w; i++) pd[i] = (uint8_t)o[i];
    }
}

inline bool plane_ok_ar(const void *p, uint32_t w, uint32_t stride, uint32_t bpp)
{
    return p && aligned16(p) && (stride % 16u) == 0 && stride >= ((w * bpp + 15u) & ~15u);
}

}  // namespace mvx

using namespace mvx;

extern "C" int mvx_Magnitude_S16_S16S16(mvx_stream_t stream, uint32_t w, uint32_t h, int16_t *dst, uint32_t dstStride,
                                        const int16_t *gx, uint32_t gxStride, const int16_t *gy, uint32_t gyStride)
{
    if (w == 0 || h == 0) return MVX_SUCCESS;
    if (!plane_ok_ar(dst, w, dstStride, 2) || !plane_ok_ar(gx, w, gxStride, 2) || !plane_ok_ar(gy, w, gyStride, 2))
        return MVX_ERROR_INVALID_PARAMETERS;
    dim3 grid(((w + 7) / 8 + AR_BLOCK - 1) / AR_BLOCK, h);
    hipLaunchKernelGGL(k_magnitude, grid, dim3(AR_BLOCK), 0, stream_of(stream), w, h, dst, dstStride, gx, gxStride, gy, gyStride);
    MVX_LAUNCH_CHECK();
    return MVX_SUCCESS;
}

extern "C" int mvx_Phase_U8_S16S16(mvx_stream_t stream, uint32_t w, uint32_t h, uint8_t *dst, uint32_t dstStride,
                                   const int16_t *gx, uint32_t gxStride, const int16_t *gy, uint32_t gyStride)
{
    if (w == 0 || h == 0) return MVX_SUCCESS;
    if (!plane_ok_ar(dst, w, dstStride, 1) || !plane_ok_ar(gx, w, gxStride, 2) || !plane_ok_ar(gy, w, gyStride, 2))
        return MVX_ERROR_INVALID_PARAMETERS;
    if ((dstStride & 7u) != 0) return MVX_ERROR_INVALID_PARAMETERS;
    dim3 grid(((w + 7) / 8 + AR_BLOCK - 1) / AR_BLOCK, h);
    hipLaunchKernelGGL(k_phase, grid, dim3(AR_BLOCK), 0, stream_of(stream), w, h, dst, dstStride, gx, gxStride, gy, gyStride);
    MVX_LAUNCH_CHECK();
    return MVX_SUCCESS;
}
