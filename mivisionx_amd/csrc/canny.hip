// canny.hip -- CannyEdgeDetector (gradient 3, L1 norm) for gfx950: Sobel -> non-max suppression with
// double threshold -> hysteresis.
//
// The reference traces hysteresis with a sequential LIFO stack on the CPU (its HIP backend has no
// trace kernel).  The traced image only depends on connectivity: a weak pixel ends at 255 iff it is
// 8-connected through weak pixels to a strong one.  Here that set is computed with a lock-free
// union-find over the edge pixels (atomicMin linking, then one find per pixel), three passes with no
// host round trip, in the 4-byte-per-pixel workspace the reference reserves for its XY stack.
#include "mvx_common.hpp"

namespace mvx {

inline bool plane_ok_c(const void *p, uint32_t w, uint32_t stride, uint32_t bpp)
{
    return p && aligned16(p) && (stride % 16u) == 0 && stride >= ((w * bpp + 15u) & ~15u);
}

// HafCpu_FastAtan2_Canny (ago_haf_cpu_arithmetic.cpp:9638-9656); statement order kept (0,0 -> 2)
MVX_DEV uint32_t canny_sector(int gx, int gy)
{
    const uint32_t ax = (uint32_t)abs(gx) & 0xffffu, ay = (uint32_t)abs(gy) & 0xffffu;
    const float d1 = (float)ax * 0.4142135623730950488016887242097f;
    const float d2 = (float)ax * 2.4142135623730950488016887242097f;
    uint32_t ret = (gx * gy) < 0 ? 3u : 1u;
    if ((float)ay <= d1) ret = 0;
    if ((float)ay >= d2) ret = 2;
    return ret;
}

MVX_DEV uint16_t canny_pack(int gx, int gyn)
{
    int16_t t = (int16_t)(abs((int)(int16_t)gx) + abs((int)(int16_t)gyn));
    t = (int16_t)(t << 2);
    t = (int16_t)(t | (int16_t)(canny_sector((int16_t)gx, (int16_t)gyn) & 3u));
    return (uint16_t)t;
}

// linear-address Sobel at (x, y) -> packed U16 (the reference's scalar loop, ago_haf_cpu_canny.cpp:120-131)
MVX_DEV uint16_t canny_flat(const uint8_t *src, long n, long stride, int x, int y)
{
    const long c0 = (long)(y - 1) * stride + x, c1 = (long)y * stride + x, c2 = (long)(y + 1) * stride + x;
    const int gx = ((int)ld_u8(src, n, c0 + 1) - (int)ld_u8(src, n, c0 - 1)) + ((int)ld_u8(src, n, c2 + 1) - (int)ld_u8(src, n, c2 - 1))
                 + 2 * ((int)ld_u8(src, n, c1 + 1) - (int)ld_u8(src, n, c1 - 1));
    const int gy = (int)ld_u8(src, n, c2 - 1) + (int)ld_u8(src, n, c2 + 1) - (int)ld_u8(src, n, c0 - 1) - (int)ld_u8(src, n, c0 + 1)
                 + 2 * ((int)ld_u8(src, n, c2) - (int)ld_u8(src, n, c0));
    return canny_pack((int16_t)gx, (int16_t)(-(int16_t)gy));
}

// CannySobel_U16_U8_3x3_L1NORM (ago_haf_cpu_canny.cpp:44-136), rows 1..h-2, widths >= 16.
// Column 0 takes the SSE path's staged values: r0[-1] = 0 and r1[-1] = r0[ALIGN16(w)-1], i.e. this
// row's vertical sum t0(w-1) when w is a multiple of 16, else 0.  Every other column equals the
// linear-address Sobel (interior SSE columns read only in-row neighbours; the scalar loop redoes
// columns aligned-1 .. w-1 with linear reads).
__global__ void __launch_bounds__(256)
k_canny_sobel(uint32_t w, uint32_t h, uint16_t *__restrict__ dst, uint32_t dstStride, const uint8_t *__restrict__ src,
              uint32_t srcStride)
{
    const uint32_t x0 = (blockIdx.x * 256 + threadIdx.x) * 8;
    const int y = 1 + (int)blockIdx.y;
    if (x0 >= w || y >= (int)h - 1) return;
    const long n = (long)h * srcStride;
    uint16_t o[8];
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const int x = (int)x0 + i;
        if (x >= (int)w) { o[i] = 0; continue; }
        if (x == 0) {
            const long bm = (long)(y - 1) * srcStride, b0 = (long)y * srcStride, bp = (long)(y + 1) * srcStride;
            const int t0_1 = (int)src[bm + 1] + (int)src[bp + 1] + 2 * (int)src[b0 + 1];
            const int t1_0 = (int)src[bp] - (int)src[bm];
            const int t1_1 = (int)src[bp + 1] - (int)src[bm + 1];
            const int t1_m1 = (w % 16 == 0) ? ((int)src[bm + w - 1] + (int)src[bp + w - 1] + 2 * (int)src[b0 + w - 1]) : 0;
            o[i] = canny_pack(t0_1, -(t1_m1 + 2 * t1_0 + t1_1));
        } else {
            o[i] = canny_flat(src, n, srcStride, x, y);
        }
    }
    uint16_t *d = (uint16_t *)((uint8_t *)dst + (long)y * dstStride) + x0;
    if (x0 + 8 <= w) {
        uint4 v;
        v.x = (uint32_t)o[0] | ((uint32_t)o[1] << 16);
        v.y = (uint32_t)o[2] | ((uint32_t)o[3] << 16);
        v.z = (uint32_t)o[4] | ((uint32_t)o[5] << 16);
        v.w = (uint32_t)o[6] | ((uint32_t)o[7] << 16);
        *reinterpret_cast<uint4 *>(d) = v;
    } else {
        for (uint32_t i = 0; x0 + i < w; i++) d[i] = o[i];
    }
}

// Widths below 16: the reference runs only its scalar loop, from x = -1 (it writes the U16 slot just
// before each row, i.e. the last slot of the previous row's stride), rows in order.  One thread
// replays that order exactly; these planes are at most 15 pixels wide.
__global__ void k_canny_sobel_narrow(uint32_t w, uint32_t h, uint16_t *dst, uint32_t dstStride, const uint8_t *src,
                                     uint32_t srcStride)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const long n = (long)h * srcStride;
    for (int y = 1; y < (int)h - 1; y++)
        for (int x = -1; x < (int)w; x++)
            *(uint16_t *)((uint8_t *)dst + (long)y * dstStride + 2L * x) = canny_flat(src, n, srcStride, x, y);
}

// CannySuppThreshold_U8XY_U16_3x3 (ago_haf_cpu_canny.cpp:549-596): x in [1, w-2], y in [1, h-2].
__global__ void __launch_bounds__(256)
k_canny_supp(uint32_t w, uint32_t h, uint8_t *__restrict__ dst, uint32_t dstStride, const uint16_t *__restrict__ mag,
             uint32_t magStride, uint32_t lower, uint32_t upper)
{
    const int x = 1 + (int)(blockIdx.x * 256 + threadIdx.x);
    const int y = 1 + (int)blockIdx.y;
    if (x >= (int)w - 1 || y >= (int)h - 1) return;
    const long ms = magStride / 2;
    const uint16_t *p = mag + (long)y * ms + x;
    const uint32_t v = p[0];
    const int m = (int)(v >> 2), a = (int)(v & 3u);
    // neighbour pairs by sector (ago_haf_cpu_canny.cpp:26-31)
    const int dx0 = (a == 0) ? -1 : (a == 1) ? 1 : (a == 2) ? 0 : -1;
    const int dy0 = (a == 0) ? 0 : (a == 1) ? -1 : (a == 2) ? -1 : -1;
    const int n0 = (int)(p[dy0 * ms + dx0] >> 2);
    const int n1 = (int)(p[-dy0 * ms - dx0] >> 2);
    const int edge = (m > n0 && m > n1) ? m : 0;
    const uint8_t out = edge > (int)upper ? 255 : (edge <= (int)lower ? 0 : 127);
    dst[(long)y * dstStride + x] = out;
}

// seed the union-find: every interior edge pixel (127 or 255) is its own root.  Labels always point
// to a smaller or equal index of an edge pixel, so finds terminate whatever the image holds.
__global__ void __launch_bounds__(256)
k_canny_init(uint32_t w, uint32_t h, const uint8_t *__restrict__ img, uint32_t stride, uint32_t *__restrict__ lab)
{
    const int x = 1 + (int)(blockIdx.x * 256 + threadIdx.x);
    const int y = 1 + (int)blockIdx.y;
    if (x >= (int)w - 1 || y >= (int)h - 1) return;
    if (img[(long)y * stride + x]) lab[(long)y * w + x] = (uint32_t)((long)y * w + x);
}

MVX_DEV uint32_t uf_find(const uint32_t *lab, uint32_t p)
{
    uint32_t l = lab[p] & 0x7fffffffu;
    while (l != p) {
        p = l;
        l = lab[p] & 0x7fffffffu;
    }
    return p;
}

// link the trees of a and b: the larger root points at the smaller (atomicMin), retried until the
// roots agree; lock-free and order independent
MVX_DEV void uf_union(uint32_t *lab, uint32_t a, uint32_t b)
{
    while (true) {
        a = uf_find(lab, a);
        b = uf_find(lab, b);
        if (a == b) return;
        if (a < b) { uint32_t t = a; a = b; b = t; }   // a > b: hang a under b
        const uint32_t old = atomicMin(&lab[a], b);
        if (old == a) return;
        a = old;
    }
}

// merge each edge pixel with its E, SW, S, SE edge neighbours (covers all 8-neighbour pairs)
__global__ void __launch_bounds__(256)
k_canny_merge(uint32_t w, uint32_t h, const uint8_t *__restrict__ img, uint32_t stride, uint32_t *__restrict__ lab)
{
    const int x = 1 + (int)(blockIdx.x * 256 + threadIdx.x);
    const int y = 1 + (int)blockIdx.y;
    if (x >= (int)w - 1 || y >= (int)h - 1) return;
    const uint8_t *r0 = img + (long)y * stride;
    if (!r0[x]) return;
    const uint8_t *r1 = r0 + stride;
    const uint32_t p = (uint32_t)((long)y * w + x);
    // only interior neighbours carry labels (the border rows/columns are outside the supp range)
    const bool right = x + 1 < (int)w - 1, below = y + 1 < (int)h - 1, left = x - 1 >= 1;
    if (right && r0[x + 1]) uf_union(lab, p, p + 1);
    if (below && left && r1[x - 1]) uf_union(lab, p, p + w - 1);
    if (below && r1[x]) uf_union(lab, p, p + w);
    if (below && right && r1[x + 1]) uf_union(lab, p, p + w + 1);
}

__global__ void __launch_bounds__(256)
k_canny_mark(uint32_t w, uint32_t h, const uint8_t *__restrict__ img, uint32_t stride, uint32_t *__restrict__ lab)
{
    const int x = 1 + (int)(blockIdx.x * 256 + threadIdx.x);
    const int y = 1 + (int)blockIdx.y;
    if (x >= (int)w - 1 || y >= (int)h - 1) return;
    if (img[(long)y * stride + x] != 255) return;
    const uint32_t r = uf_find(lab, (uint32_t)((long)y * w + x));
    atomicOr(&lab[r], 0x80000000u);
}

__global__ void __launch_bounds__(256)
k_canny_final(uint32_t w, uint32_t h, uint8_t *__restrict__ img, uint32_t stride, const uint32_t *__restrict__ lab)
{
    const int x = 1 + (int)(blockIdx.x * 256 + threadIdx.x);
    const int y = 1 + (int)blockIdx.y;
    if (x >= (int)w - 1 || y >= (int)h - 1) return;
    uint8_t *p = img + (long)y * stride + x;
    if (*p != 127) return;
    const uint32_t r = uf_find(lab, (uint32_t)((long)y * w + x));
    *p = (lab[r] & 0x80000000u) ? 255 : 0;
}

}  // namespace mvx

using namespace mvx;

extern "C" int mvx_CannySobel_U16_U8_3x3_L1NORM(mvx_stream_t stream, uint32_t w, uint32_t h, uint16_t *dst,
                                                uint32_t dstStride, const uint8_t *src, uint32_t srcStride)
{
    if (w == 0 || h == 0) return MVX_SUCCESS;
    if (!plane_ok_c(dst, w, dstStride, 2) || !plane_ok_c(src, w, srcStride, 1)) return MVX_ERROR_INVALID_PARAMETERS;
    if (h < 3) return MVX_SUCCESS;
    if (w < 16) {
        hipLaunchKernelGGL(k_canny_sobel_narrow, dim3(1), dim3(64), 0, stream_of(stream), w, h, dst, dstStride, src, srcStride);
    } else {
        dim3 grid(((w + 7) / 8 + 255) / 256, h - 2);
        hipLaunchKernelGGL(k_canny_sobel, grid, dim3(256), 0, stream_of(stream), w, h, dst, dstStride, src, srcStride);
    }
    MVX_LAUNCH_CHECK();
    return MVX_SUCCESS;
}

extern "C" int mvx_CannySuppThreshold_U8XY_U16_3x3(mvx_stream_t stream, uint32_t w, uint32_t h, uint8_t *dst,
                                                   uint32_t dstStride, const uint16_t *mag, uint32_t magStride,
                                                   void *xyStack, uint32_t capacityOfXY, uint32_t lower, uint32_t upper)
{
    if (w == 0 || h == 0) return MVX_SUCCESS;
    if (!plane_ok_c(dst, w, dstStride, 1) || !plane_ok_c(mag, w, magStride, 2) || !xyStack ||
        (uint64_t)capacityOfXY < (uint64_t)w * h)
        return MVX_ERROR_INVALID_PARAMETERS;
    if (w < 3 || h < 3) return MVX_SUCCESS;
    dim3 grid((w - 2 + 255) / 256, h - 2);
    hipLaunchKernelGGL(k_canny_supp, grid, dim3(256), 0, stream_of(stream), w, h, dst, dstStride, mag, magStride,
                       lower, upper);
    MVX_LAUNCH_CHECK();
    return MVX_SUCCESS;
}

extern "C" int mvx_CannyEdgeTrace_U8_U8XY(mvx_stream_t stream, uint32_t w, uint32_t h, uint8_t *img, uint32_t stride,
                                          void *xyStack, uint32_t capacityOfXY)
{
    if (w == 0 || h == 0) return MVX_SUCCESS;
    if (!plane_ok_c(img, w, stride, 1) || !xyStack || (uint64_t)capacityOfXY < (uint64_t)w * h)
        return MVX_ERROR_INVALID_PARAMETERS;
    if (w < 3 || h < 3) return MVX_SUCCESS;
    dim3 grid((w - 2 + 255) / 256, h - 2);
    uint32_t *lab = (uint32_t *)xyStack;
    hipLaunchKernelGGL(k_canny_init, grid, dim3(256), 0, stream_of(stream), w, h, img, stride, lab);
    hipLaunchKernelGGL(k_canny_merge, grid, dim3(256), 0, stream_of(stream), w, h, img, stride, lab);
    hipLaunchKernelGGL(k_canny_mark, grid, dim3(256), 0, stream_of(stream), w, h, img, stride, lab);
    hipLaunchKernelGGL(k_canny_final, grid, dim3(256), 0, stream_of(stream), w, h, img, stride, lab);
    MVX_LAUNCH_CHECK();
    return MVX_SUCCESS;
}
