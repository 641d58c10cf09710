// color.hip -- the ColorConvert families for gfx950 beyond RGB <-> IYUV / NV12 -> RGB (geometric.hip):
// YUV (UYVY, YUYV, IYUV, NV12, NV21) -> RGB / RGBX, RGB / RGBX -> NV12 / IYUV / YUV4, RGB <-> RGBX and
// the byte-moving YUV -> YUV conversions, each reproducing the reference CPU kernel it replaces bit
// for bit (ago_haf_cpu_color_convert.cpp; the SSE bodies and the scalar tails differ in rounding or
// association in places, and the column split between them is reproduced).
//
// One thread per 2x2 pixel block for 4:2:0 data (the block shares a chroma sample), per horizontal
// pixel pair for 4:2:2 data, per pixel otherwise.  The kernels are HBM-bound byte shuffles with a few
// float operations per pixel; all float sequences follow the reference's order without contraction.
#include "mvx_common.hpp"

namespace mvx {

namespace {

enum Yuv { F_IYUV = 0, F_NV12 = 1, F_NV21 = 2, F_UYVY = 3, F_YUYV = 4 };

MVX_DEV uint32_t clamp_trunc(float f) { return (uint32_t)fminf(fmaxf(f, 0.0f), 255.0f); }
MVX_DEV int sat(int v, int hi) { return min(max(v, 0), hi); }
// (vx_uint8) of a double as the reference's x86 build converts it: to int32, then the low byte
MVX_DEV uint32_t u8_of(double d) { return (uint32_t)(int)d & 0xffu; }

// one RGB(X) pixel from Y, U, V (ago_haf_cpu_color_convert.cpp:280-1732); tail422: the 4:2:2 scalar
// tail's (Y - u*0.1873) - v*0.4681 grouping of G
MVX_DEV void yuv_pixel(uint32_t Y, uint32_t U, uint32_t V, bool tail422, uint32_t &r, uint32_t &g, uint32_t &b)
{
    const float fy = (float)Y, u = (float)U - 128.0f, v = (float)V - 128.0f;
    const float gg = tail422 ? (fy - u * 0.1873f) - v * 0.4681f : fy - ((u * 0.1873f) + (v * 0.4681f));
    r = clamp_trunc(fy + v * 1.5748f);
    g = clamp_trunc(gg);
    b = clamp_trunc(fy + u * 1.8556f);
}

struct Planes3 {
    const uint8_t *p0, *p1, *p2;
    uint32_t s0, s1, s2;
};

// YUV -> RGB / RGBX: thread = a horizontal pixel pair (x = 2i) of row y
template <int FMT, int OUT>
__global__ void __launch_bounds__(256)
k_rgb_from_yuv(uint32_t w, uint32_t h, uint8_t *__restrict__ dst, uint32_t dstStride, Planes3 s)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
    const uint32_t x = 2 * i;
    if (x >= w || y >= h) return;
    uint32_t Y0, Y1, U, V;
    if (FMT == F_UYVY || FMT == F_YUYV) {
        const uint32_t q = *reinterpret_cast<const uint32_t *>(s.p0 + (uint64_t)y * s.s0 + 2ull * x);
        const uint32_t b0 = q & 0xffu, b1 = (q >> 8) & 0xffu, b2 = (q >> 16) & 0xffu, b3 = q >> 24;
        if (FMT == F_UYVY) { U = b0; Y0 = b1; V = b2; Y1 = b3; }
        else { Y0 = b0; U = b1; Y1 = b2; V = b3; }
    } else {
        const uint8_t *yr = s.p0 + (uint64_t)y * s.s0 + x;
        Y0 = yr[0];
        Y1 = yr[1];
        const uint64_t c = (uint64_t)(y >> 1) * s.s1;
        if (FMT == F_IYUV) {
            U = s.p1[c + i];
            V = s.p2[(uint64_t)(y >> 1) * s.s2 + i];
        } else {
            const uint32_t a = s.p1[c + x], bb = s.p1[c + x + 1];
            U = FMT == F_NV12 ? a : bb;
            V = FMT == F_NV12 ? bb : a;
        }
    }
    const bool tail = (FMT == F_UYVY || FMT == F_YUYV) && x >= (w & ~7u);
    uint32_t r0, g0, b0, r1, g1, b1;
    yuv_pixel(Y0, U, V, tail, r0, g0, b0);
    yuv_pixel(Y1, U, V, tail, r1, g1, b1);
    uint8_t *d = dst + (uint64_t)y * dstStride + (uint64_t)OUT * x;
    if (OUT == 4) {
        const uint32_t lo = r0 | (g0 << 8) | (b0 << 16) | 0xff000000u, hi = r1 | (g1 << 8) | (b1 << 16) | 0xff000000u;
        *reinterpret_cast<uint2 *>(d) = make_uint2(lo, hi);
    } else {
        uint16_t *d16 = reinterpret_cast<uint16_t *>(d);   // 6x is even: 2-byte aligned
        d16[0] = (uint16_t)(r0 | (g0 << 8));
        d16[1] = (uint16_t)(b0 | (r1 << 8));
        d16[2] = (uint16_t)(g1 | (b1 << 8));
    }
}

// RGB / RGBX -> IYUV / NV12 (HafCpu_ColorConvert_IYUV_RGB, _NV12_RGB, _IYUV_RGBX, _NV12_RGBX): thread =
// the 2x2 block at (2i, 2j); columns below w & ~3 follow the SSE body, the rest the scalar tail
template <int INBPP, bool NV12>
__global__ void __launch_bounds__(256)
k_yuv420_from_rgb(uint32_t w, uint32_t h, uint8_t *__restrict__ yp, uint32_t ys, uint8_t *__restrict__ up, uint32_t us,
                  uint8_t *__restrict__ vp, uint32_t vs, const uint8_t *__restrict__ src, uint32_t ss)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x, j = blockIdx.y;
    const uint32_t x = 2 * i, r = 2 * j;
    if (x >= w || r >= h) return;
    float R[4], G[4], B[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const uint8_t *p = src + (uint64_t)(r + (k >> 1)) * ss + (uint64_t)INBPP * (x + (k & 1));
        R[k] = (float)p[0];
        G[k] = (float)p[1];
        B[k] = (float)p[2];
    }
    uint32_t yv[4], uo, vo;
    if (x < (w & ~3u)) {
        int uu[4], vv[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const float yf = ((R[k] * 0.2126f) + (G[k] * 0.7152f)) + (B[k] * 0.0722f);
            const float uf = ((R[k] * -0.1146f) + (G[k] * -0.3854f)) + (B[k] * 0.5f);
            const float vf = ((R[k] * 0.5f) + (G[k] * -0.4542f)) + (B[k] * -0.0458f);
            if (INBPP == 3) {
                yv[k] = (uint32_t)sat((int)yf, 255);
                uu[k] = sat((int)(uf + 128.0f), 65535);
                vv[k] = sat((int)(vf + 128.0f), 65535);
            } else {
                yv[k] = (uint32_t)sat(__float2int_rn(yf), 255);
                uu[k] = sat(__float2int_rn(uf) + 128, 65535);
                vv[k] = sat(__float2int_rn(vf) + 128, 65535);
            }
        }
        const int rnd = INBPP == 3 ? 1 : 0;
        uo = (uint32_t)sat((((uu[0] + uu[2] + 1) >> 1) + ((uu[1] + uu[3] + 1) >> 1) + rnd) >> 1, 255);
        vo = (uint32_t)sat((((vv[0] + vv[2] + 1) >> 1) + ((vv[1] + vv[3] + 1) >> 1) + rnd) >> 1, 255);
    } else {
        float U = 0.0f, V = 0.0f;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const double yd = (double)((R[k] * 0.2126f) + (G[k] * 0.7152f)) + (double)B[k] * 0.0722;
            yv[k] = INBPP == 3 ? u8_of(yd) : u8_of(round(yd));
            const float uf = ((R[k] * -0.1146f) + (G[k] * -0.3854f)) + (B[k] * 0.5f) + 128.0f;
            const float vf = ((R[k] * 0.5f) + (G[k] * -0.4542f)) + (B[k] * -0.0458f) + 128.0f;
            U = k ? U + uf : uf;
            V = k ? V + vf : vf;
        }
        U = U / 4.0f;
        V = V / 4.0f;
        uo = INBPP == 3 ? u8_of((double)U) : u8_of(round((double)U));
        vo = INBPP == 3 ? u8_of((double)V) : u8_of(round((double)V));
    }
    *reinterpret_cast<uint16_t *>(yp + (uint64_t)r * ys + x) = (uint16_t)(yv[0] | (yv[1] << 8));
    *reinterpret_cast<uint16_t *>(yp + (uint64_t)(r + 1) * ys + x) = (uint16_t)(yv[2] | (yv[3] << 8));
    if (NV12) {
        *reinterpret_cast<uint16_t *>(up + (uint64_t)j * us + x) = (uint16_t)(uo | (vo << 8));
    } else {
        up[(uint64_t)j * us + i] = (uint8_t)uo;
        vp[(uint64_t)j * vs + i] = (uint8_t)vo;
    }
}

// RGB / RGBX -> YUV4 (HafCpu_ColorConvert_YUV4_RGB / _YUV4_RGBX): thread = one pixel; columns below
// w & ~15 round to nearest even (SSE body), the rest through round() in double (scalar tail; U's green
// weight is a double constant there)
template <int INBPP>
__global__ void __launch_bounds__(256)
k_yuv4_from_rgb(uint32_t w, uint32_t h, uint8_t *__restrict__ yp, uint32_t ys, uint8_t *__restrict__ up, uint32_t us,
                uint8_t *__restrict__ vp, uint32_t vs, const uint8_t *__restrict__ src, uint32_t ss)
{
    const uint32_t x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
    if (x >= w || y >= h) return;
    const uint8_t *p = src + (uint64_t)y * ss + (uint64_t)INBPP * x;
    const float R = (float)p[0], G = (float)p[1], B = (float)p[2];
    uint32_t Y, U, V;
    if (x < (w & ~15u)) {
        Y = (uint32_t)sat(__float2int_rn(((R * 0.2126f) + (G * 0.7152f)) + (B * 0.0722f)), 255);
        U = (uint32_t)sat(__float2int_rn(((R * -0.1146f) + (G * -0.3854f)) + (B * 0.5f)) + 128, 255);
        V = (uint32_t)sat(__float2int_rn(((R * 0.5f) + (G * -0.4542f)) + (B * -0.0458f)) + 128, 255);
    } else {
        Y = u8_of(round((double)((R * 0.2126f) + (G * 0.7152f)) + (double)B * 0.0722));
        U = u8_of(round((((double)(R * -0.1146f) + (double)G * -0.3854) + (double)(B * 0.5f)) + 128.0));
        V = u8_of(round((double)((((R * 0.5f) + (G * -0.4542f)) + (B * -0.0458f)) + 128.0f)));
    }
    yp[(uint64_t)y * ys + x] = (uint8_t)Y;
    up[(uint64_t)y * us + x] = (uint8_t)U;
    vp[(uint64_t)y * vs + x] = (uint8_t)V;
}

// RGBX -> luma alone (HafCpu_ColorConvert_Y_RGBX, the kernel kept when only Y is read): columns below
// w & ~15 truncate the float sum (SSE body), the rest a double sum with a double blue weight (scalar tail)
__global__ void __launch_bounds__(256)
k_y_rgbx(uint32_t w, uint32_t h, uint8_t *__restrict__ dst, uint32_t ds, const uint8_t *__restrict__ src, uint32_t ss)
{
    const uint32_t x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
    if (x >= w || y >= h) return;
    const uint32_t q = *reinterpret_cast<const uint32_t *>(src + (uint64_t)y * ss + 4ull * x);
    const float R = (float)(q & 0xffu), G = (float)((q >> 8) & 0xffu), B = (float)((q >> 16) & 0xffu);
    const uint32_t v = x < (w & ~15u) ? (uint32_t)sat((int)(((R * 0.2126f) + (G * 0.7152f)) + (B * 0.0722f)), 255)
                                      : u8_of((double)((R * 0.2126f) + (G * 0.7152f)) + (double)B * 0.0722);
    dst[(uint64_t)y * ds + x] = (uint8_t)v;
}

// RGB <-> RGBX (HafCpu_ColorConvert_RGBX_RGB sets X = 255, _RGB_RGBX drops it): thread = one pixel
template <bool TO_RGBX>
__global__ void __launch_bounds__(256)
k_rgb_rgbx(uint32_t w, uint32_t h, uint8_t *__restrict__ dst, uint32_t ds, const uint8_t *__restrict__ src, uint32_t ss)
{
    const uint32_t x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
    if (x >= w || y >= h) return;
    const uint8_t *s = src + (uint64_t)y * ss + (TO_RGBX ? 3ull : 4ull) * x;
    uint8_t *d = dst + (uint64_t)y * ds + (TO_RGBX ? 4ull : 3ull) * x;
    const uint8_t r = s[0], g = s[1], b = s[2];
    if (TO_RGBX) {
        *reinterpret_cast<uint32_t *>(d) = (uint32_t)r | ((uint32_t)g << 8) | ((uint32_t)b << 16) | 0xff000000u;
    } else {
        d[0] = r; d[1] = g; d[2] = b;
    }
}

// UYVY / YUYV -> IYUV / NV12 (HafCpu_FormatConvert_IYUV_UYVY, _IYUV_YUYV, _NV12_UYVY, _NV12_YUYV): thread =
// the 2x2 block at (2i, 2j); luma copied, chroma the floor average of the two rows' samples
template <int FMT, bool NV12>
__global__ void __launch_bounds__(256)
k_yuv420_from_422(uint32_t w, uint32_t h, uint8_t *__restrict__ yp, uint32_t ys, uint8_t *__restrict__ up, uint32_t us,
                  uint8_t *__restrict__ vp, uint32_t vs, const uint8_t *__restrict__ src, uint32_t ss)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x, j = blockIdx.y;
    const uint32_t x = 2 * i, r = 2 * j;
    if (x >= w || r >= h) return;
    const uint32_t q0 = *reinterpret_cast<const uint32_t *>(src + (uint64_t)r * ss + 2ull * x);
    const uint32_t q1 = *reinterpret_cast<const uint32_t *>(src + (uint64_t)(r + 1) * ss + 2ull * x);
    const int yo = FMT == F_UYVY ? 8 : 0, uo = FMT == F_UYVY ? 0 : 8, vo = FMT == F_UYVY ? 16 : 24;
    const uint32_t y00 = (q0 >> yo) & 0xffu, y01 = (q0 >> (yo + 16)) & 0xffu, y10 = (q1 >> yo) & 0xffu, y11 = (q1 >> (yo + 16)) & 0xffu;
    const uint32_t u = (((q0 >> uo) & 0xffu) + ((q1 >> uo) & 0xffu)) >> 1, v = (((q0 >> vo) & 0xffu) + ((q1 >> vo) & 0xffu)) >> 1;
    *reinterpret_cast<uint16_t *>(yp + (uint64_t)r * ys + x) = (uint16_t)(y00 | (y01 << 8));
    *reinterpret_cast<uint16_t *>(yp + (uint64_t)(r + 1) * ys + x) = (uint16_t)(y10 | (y11 << 8));
    if (NV12) {
        *reinterpret_cast<uint16_t *>(up + (uint64_t)j * us + x) = (uint16_t)(u | (v << 8));
    } else {
        up[(uint64_t)j * us + i] = (uint8_t)u;
        vp[(uint64_t)j * vs + i] = (uint8_t)v;
    }
}

// interleaved UV (cw x ch samples) -> separate U and V (HafCpu_FormatConvert_IUV_UV12); UP2: each sample
// to a 2x2 block of full-size planes (HafCpu_FormatConvert_UV_UV12).  Thread = one sample.
template <bool UP2>
__global__ void __launch_bounds__(256)
k_split_uv(uint32_t cw, uint32_t ch, uint8_t *__restrict__ up, uint32_t us, uint8_t *__restrict__ vp, uint32_t vs,
           const uint8_t *__restrict__ src, uint32_t ss)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x, j = blockIdx.y;
    if (i >= cw || j >= ch) return;
    const uint16_t uv = *reinterpret_cast<const uint16_t *>(src + (uint64_t)j * ss + 2ull * i);
    const uint8_t u = (uint8_t)(uv & 0xffu), v = (uint8_t)(uv >> 8);
    if (UP2) {
        const uint16_t uu = (uint16_t)(u | (u << 8)), vv = (uint16_t)(v | (v << 8));
        *reinterpret_cast<uint16_t *>(up + (uint64_t)(2 * j) * us + 2 * i) = uu;
        *reinterpret_cast<uint16_t *>(up + (uint64_t)(2 * j + 1) * us + 2 * i) = uu;
        *reinterpret_cast<uint16_t *>(vp + (uint64_t)(2 * j) * vs + 2 * i) = vv;
        *reinterpret_cast<uint16_t *>(vp + (uint64_t)(2 * j + 1) * vs + 2 * i) = vv;
    } else {
        up[(uint64_t)j * us + i] = u;
        vp[(uint64_t)j * vs + i] = v;
    }
}

// separate U, V (cw x ch) -> interleaved UV (HafCpu_FormatConvert_UV12_IUV); thread = one sample
__global__ void __launch_bounds__(256)
k_merge_uv(uint32_t cw, uint32_t ch, uint8_t *__restrict__ dst, uint32_t ds, const uint8_t *__restrict__ up, uint32_t us,
           const uint8_t *__restrict__ vp, uint32_t vs)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x, j = blockIdx.y;
    if (i >= cw || j >= ch) return;
    *reinterpret_cast<uint16_t *>(dst + (uint64_t)j * ds + 2ull * i) =
        (uint16_t)(up[(uint64_t)j * us + i] | (vp[(uint64_t)j * vs + i] << 8));
}

// each source pixel to a 2x2 block (HafCpu_ScaleUp2x2_U8_U8); thread = one source pixel of the
// dstWidth/2 x dstHeight/2 source
__global__ void __launch_bounds__(256)
k_scale_up2x2(uint32_t dw, uint32_t dh, uint8_t *__restrict__ dst, uint32_t ds, const uint8_t *__restrict__ src, uint32_t ss)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x, j = blockIdx.y;
    if (2 * i >= dw || 2 * j >= dh) return;
    const uint8_t p = src[(uint64_t)j * ss + i];
    const uint16_t pp = (uint16_t)(p | (p << 8));
    *reinterpret_cast<uint16_t *>(dst + (uint64_t)(2 * j) * ds + 2 * i) = pp;
    *reinterpret_cast<uint16_t *>(dst + (uint64_t)(2 * j + 1) * ds + 2 * i) = pp;
}

// device plane check: base 16-byte aligned, stride a multiple of 16 covering `rowBytes`
bool plane_ok(const void *p, uint32_t stride, uint64_t rowBytes)
{
    return p && aligned16(p) && (stride % 16u) == 0 && stride >= rowBytes;
}

dim3 grid_of(uint32_t threadsX, uint32_t rows) { return dim3((threadsX + 255) / 256, rows); }

}  // namespace
}  // namespace mvx

using namespace mvx;

// ---------------------------------------------------------------- YUV -> RGB / RGBX ----------------
#define MVX_RGB_FROM_422(NAME, FMT, OUT)                                                                                       \
    extern "C" int NAME(mvx_stream_t stream, uint32_t w, uint32_t h, uint8_t *dst, uint32_t dstStride, const uint8_t *src,  \
                        uint32_t srcStride)                                                                                    \
    {                                                                                                                          \
        if (!w || !h) return MVX_SUCCESS;                                                                                      \
        if ((w & 1u) || !plane_ok(dst, dstStride, (uint64_t)OUT * w) || !plane_ok(src, srcStride, 2ull * w))                   \
            return MVX_ERROR_INVALID_PARAMETERS;                                                                               \
        hipLaunchKernelGGL((k_rgb_from_yuv<FMT, OUT>), grid_of(w / 2, h), dim3(256), 0, stream_of(stream), w, h, dst, dstStride, \
                           Planes3{src, nullptr, nullptr, srcStride, 0u, 0u});                                                 \
        MVX_LAUNCH_CHECK();                                                                                                    \
        return MVX_SUCCESS;                                                                                                    \
    }
MVX_RGB_FROM_422(mvx_ColorConvert_RGB_UYVY, F_UYVY, 3)
MVX_RGB_FROM_422(mvx_ColorConvert_RGB_YUYV, F_YUYV, 3)
MVX_RGB_FROM_422(mvx_ColorConvert_RGBX_UYVY, F_UYVY, 4)
MVX_RGB_FROM_422(mvx_ColorConvert_RGBX_YUYV, F_YUYV, 4)
#undef MVX_RGB_FROM_422

#define MVX_RGB_FROM_SEMI(NAME, FMT, OUT)                                                                                      \
    extern "C" int NAME(mvx_stream_t stream, uint32_t w, uint32_t h, uint8_t *dst, uint32_t dstStride, const uint8_t *luma,  \
                        uint32_t lumaStride, const uint8_t *chroma, uint32_t chromaStride)                                     \
    {                                                                                                                          \
        if (!w || !h) return MVX_SUCCESS;                                                                                      \
        if ((w & 1u) || (h & 1u) || !plane_ok(dst, dstStride, (uint64_t)OUT * w) || !plane_ok(luma, lumaStride, w) ||          \
            !plane_ok(chroma, chromaStride, w))                                                                                \
            return MVX_ERROR_INVALID_PARAMETERS;                                                                               \
        hipLaunchKernelGGL((k_rgb_from_yuv<FMT, OUT>), grid_of(w / 2, h), dim3(256), 0, stream_of(stream), w, h, dst, dstStride, \
                           Planes3{luma, chroma, nullptr, lumaStride, chromaStride, 0u});                                      \
        MVX_LAUNCH_CHECK();                                                                                                    \
        return MVX_SUCCESS;                                                                                                    \
    }
MVX_RGB_FROM_SEMI(mvx_ColorConvert_RGB_NV21, F_NV21, 3)
MVX_RGB_FROM_SEMI(mvx_ColorConvert_RGBX_NV12, F_NV12, 4)
MVX_RGB_FROM_SEMI(mvx_ColorConvert_RGBX_NV21, F_NV21, 4)
#undef MVX_RGB_FROM_SEMI

extern "C" int mvx_ColorConvert_RGBX_IYUV(mvx_stream_t stream, uint32_t w, uint32_t h, uint8_t *dst, uint32_t dstStride,
                                          const uint8_t *y, uint32_t yStride, const uint8_t *u, uint32_t uStride, const uint8_t *v,
                                          uint32_t vStride)
{
    if (!w || !h) return MVX_SUCCESS;
    if ((w & 1u) || (h & 1u) || !plane_ok(dst, dstStride, 4ull * w) || !plane_ok(y, yStride, w) || !plane_ok(u, uStride, w / 2) ||
        !plane_ok(v, vStride, w / 2))
        return MVX_ERROR_INVALID_PARAMETERS;
    hipLaunchKernelGGL((k_rgb_from_yuv<F_IYUV, 4>), grid_of(w / 2, h), dim3(256), 0, stream_of(stream), w, h, dst, dstStride,
                       Planes3{y, u, v, yStride, uStride, vStride});
    MVX_LAUNCH_CHECK();
    return MVX_SUCCESS;
}

// ---------------------------------------------------------------- RGB / RGBX -> YUV ---------------
extern "C" int mvx_ColorConvert_NV12_RGB(mvx_stream_t stream, uint32_t w, uint32_t h, uint8_t *luma, uint32_t lumaStride,
                                         uint8_t *chroma, uint32_t chromaStride, const uint8_t *src, uint32_t srcStride)
{
    if (!w || !h) return MVX_SUCCESS;
    if ((w & 1u) || (h & 1u) || !plane_ok(luma, lumaStride, w) || !plane_ok(chroma, chromaStride, w) || !plane_ok(src, srcStride, 3ull * w))
        return MVX_ERROR_INVALID_PARAMETERS;
    hipLaunchKernelGGL((k_yuv420_from_rgb<3, true>), grid_of(w / 2, h / 2), dim3(256), 0, stream_of(stream), w, h, luma, lumaStride,
                       chroma, chromaStride, nullptr, 0u, src, srcStride);
    MVX_LAUNCH_CHECK();
    return MVX_SUCCESS;
}

extern "C" int mvx_ColorConvert_NV12_RGBX(mvx_stream_t stream, uint32_t w, uint32_t h, uint8_t *luma, uint32_t lumaStride,
                                          uint8_t *chroma, uint32_t chromaStride, const uint8_t *src, uint32_t srcStride)
{
    if (!w || !h) return MVX_SUCCESS;
    if ((w & 1u) || (h & 1u) || !plane_ok(luma, lumaStride, w) || !plane_ok(chroma, chromaStride, w) || !plane_ok(src, srcStride, 4ull * w))
        return MVX_ERROR_INVALID_PARAMETERS;
    hipLaunchKernelGGL((k_yuv420_from_rgb<4, true>), grid_of(w / 2, h / 2), dim3(256), 0, stream_of(stream), w, h, luma, lumaStride,
                       chroma, chromaStride, nullptr, 0u, src, srcStride);
    MVX_LAUNCH_CHECK();
    return MVX_SUCCESS;
}

extern "C" int mvx_ColorConvert_IYUV_RGBX(mvx_stream_t stream, uint32_t w, uint32_t h, uint8_t *y, uint32_t yStride, uint8_t *u,
                                          uint32_t uStride, uint8_t *v, uint32_t vStride, const uint8_t *src, uint32_t srcStride)
{
    if (!w || !h) return MVX_SUCCESS;
    if ((w & 1u) || (h & 1u) || !plane_ok(y, yStride, w) || !plane_ok(u, uStride, w / 2) || !plane_ok(v, vStride, w / 2) ||
        !plane_ok(src, srcStride, 4ull * w))
        return MVX_ERROR_INVALID_PARAMETERS;
    hipLaunchKernelGGL((k_yuv420_from_rgb<4, false>), grid_of(w / 2, h / 2), dim3(256), 0, stream_of(stream), w, h, y, yStride, u,
                       uStride, v, vStride, src, srcStride);
    MVX_LAUNCH_CHECK();
    return MVX_SUCCESS;
}

#define MVX_YUV4_FROM(NAME, INBPP)                                                                                             \
    extern "C" int NAME(mvx_stream_t stream, uint32_t w, uint32_t h, uint8_t *y, uint32_t yStride, uint8_t *u, uint32_t uStride, \
                        uint8_t *v, uint32_t vStride, const uint8_t *src, uint32_t srcStride)                                  \
    {                                                                                                                          \
        if (!w || !h) return MVX_SUCCESS;                                                                                      \
        if (!plane_ok(y, yStride, w) || !plane_ok(u, uStride, w) || !plane_ok(v, vStride, w) ||                                \
            !plane_ok(src, srcStride, (uint64_t)INBPP * w))                                                                    \
            return MVX_ERROR_INVALID_PARAMETERS;                                                                               \
        hipLaunchKernelGGL((k_yuv4_from_rgb<INBPP>), grid_of(w, h), dim3(256), 0, stream_of(stream), w, h, y, yStride, u, uStride, \
                           v, vStride, src, srcStride);                                                                        \
        MVX_LAUNCH_CHECK();                                                                                                    \
        return MVX_SUCCESS;                                                                                                    \
    }
MVX_YUV4_FROM(mvx_ColorConvert_YUV4_RGB, 3)
MVX_YUV4_FROM(mvx_ColorConvert_YUV4_RGBX, 4)
#undef MVX_YUV4_FROM

extern "C" int mvx_ColorConvert_Y_RGBX(mvx_stream_t stream, uint32_t w, uint32_t h, uint8_t *dst, uint32_t dstStride,
                                       const uint8_t *src, uint32_t srcStride)
{
    if (!w || !h) return MVX_SUCCESS;
    if (!plane_ok(dst, dstStride, w) || !plane_ok(src, srcStride, 4ull * w)) return MVX_ERROR_INVALID_PARAMETERS;
    hipLaunchKernelGGL(k_y_rgbx, grid_of(w, h), dim3(256), 0, stream_of(stream), w, h, dst, dstStride, src, srcStride);
    MVX_LAUNCH_CHECK();
    return MVX_SUCCESS;
}

// ---------------------------------------------------------------- RGB <-> RGBX ---------------------
extern "C" int mvx_ColorConvert_RGBX_RGB(mvx_stream_t stream, uint32_t w, uint32_t h, uint8_t *dst, uint32_t dstStride,
                                         const uint8_t *src, uint32_t srcStride)
{
    if (!w || !h) return MVX_SUCCESS;
    if (!plane_ok(dst, dstStride, 4ull * w) || !plane_ok(src, srcStride, 3ull * w)) return MVX_ERROR_INVALID_PARAMETERS;
    hipLaunchKernelGGL(k_rgb_rgbx<true>, grid_of(w, h), dim3(256), 0, stream_of(stream), w, h, dst, dstStride, src, srcStride);
    MVX_LAUNCH_CHECK();
    return MVX_SUCCESS;
}

extern "C" int mvx_ColorConvert_RGB_RGBX(mvx_stream_t stream, uint32_t w, uint32_t h, uint8_t *dst, uint32_t dstStride,
                                         const uint8_t *src, uint32_t srcStride)
{
    if (!w || !h) return MVX_SUCCESS;
    if (!plane_ok(dst, dstStride, 3ull * w) || !plane_ok(src, srcStride, 4ull * w)) return MVX_ERROR_INVALID_PARAMETERS;
    hipLaunchKernelGGL(k_rgb_rgbx<false>, grid_of(w, h), dim3(256), 0, stream_of(stream), w, h, dst, dstStride, src, srcStride);
    MVX_LAUNCH_CHECK();
    return MVX_SUCCESS;
}

// ---------------------------------------------------------------- YUV -> YUV -----------------------
#define MVX_420_FROM_422(NAME, FMT)                                                                                            \
    extern "C" int NAME(mvx_stream_t stream, uint32_t w, uint32_t h, uint8_t *y, uint32_t yStride, uint8_t *u, uint32_t uStride, \
                        uint8_t *v, uint32_t vStride, const uint8_t *src, uint32_t srcStride)                                  \
    {                                                                                                                          \
        if (!w || !h) return MVX_SUCCESS;                                                                                      \
        if ((w & 1u) || (h & 1u) || !plane_ok(y, yStride, w) || !plane_ok(u, uStride, w / 2) || !plane_ok(v, vStride, w / 2) || \
            !plane_ok(src, srcStride, 2ull * w))                                                                               \
            return MVX_ERROR_INVALID_PARAMETERS;                                                                               \
        hipLaunchKernelGGL((k_yuv420_from_422<FMT, false>), grid_of(w / 2, h / 2), dim3(256), 0, stream_of(stream), w, h, y,     \
                           yStride, u, uStride, v, vStride, src, srcStride);                                                   \
        MVX_LAUNCH_CHECK();                                                                                                    \
        return MVX_SUCCESS;                                                                                                    \
    }
MVX_420_FROM_422(mvx_FormatConvert_IYUV_UYVY, F_UYVY)
MVX_420_FROM_422(mvx_FormatConvert_IYUV_YUYV, F_YUYV)
#undef MVX_420_FROM_422

#define MVX_NV12_FROM_422(NAME, FMT)                                                                                           \
    extern "C" int NAME(mvx_stream_t stream, uint32_t w, uint32_t h, uint8_t *luma, uint32_t lumaStride, uint8_t *chroma,      \
                        uint32_t chromaStride, const uint8_t *src, uint32_t srcStride)                                         \
    {                                                                                                                          \
        if (!w || !h) return MVX_SUCCESS;                                                                                      \
        if ((w & 1u) || (h & 1u) || !plane_ok(luma, lumaStride, w) || !plane_ok(chroma, chromaStride, w) ||                    \
            !plane_ok(src, srcStride, 2ull * w))                                                                               \
            return MVX_ERROR_INVALID_PARAMETERS;                                                                               \
        hipLaunchKernelGGL((k_yuv420_from_422<FMT, true>), grid_of(w / 2, h / 2), dim3(256), 0, stream_of(stream), w, h, luma,   \
                           lumaStride, chroma, chromaStride, nullptr, 0u, src, srcStride);                                     \
        MVX_LAUNCH_CHECK();                                                                                                    \
        return MVX_SUCCESS;                                                                                                    \
    }
MVX_NV12_FROM_422(mvx_FormatConvert_NV12_UYVY, F_UYVY)
MVX_NV12_FROM_422(mvx_FormatConvert_NV12_YUYV, F_YUYV)
#undef MVX_NV12_FROM_422

extern "C" int mvx_FormatConvert_IUV_UV12(mvx_stream_t stream, uint32_t cw, uint32_t ch, uint8_t *u, uint32_t uStride, uint8_t *v,
                                          uint32_t vStride, const uint8_t *uv, uint32_t uvStride)
{
    if (!cw || !ch) return MVX_SUCCESS;
    if (!plane_ok(u, uStride, cw) || !plane_ok(v, vStride, cw) || !plane_ok(uv, uvStride, 2ull * cw)) return MVX_ERROR_INVALID_PARAMETERS;
    hipLaunchKernelGGL(k_split_uv<false>, grid_of(cw, ch), dim3(256), 0, stream_of(stream), cw, ch, u, uStride, v, vStride, uv, uvStride);
    MVX_LAUNCH_CHECK();
    return MVX_SUCCESS;
}

extern "C" int mvx_FormatConvert_UV_UV12(mvx_stream_t stream, uint32_t w, uint32_t h, uint8_t *u, uint32_t uStride, uint8_t *v,
                                         uint32_t vStride, const uint8_t *uv, uint32_t uvStride)
{
    if (!w || !h) return MVX_SUCCESS;
    if ((w & 1u) || (h & 1u) || !plane_ok(u, uStride, w) || !plane_ok(v, vStride, w) || !plane_ok(uv, uvStride, w))
        return MVX_ERROR_INVALID_PARAMETERS;
    hipLaunchKernelGGL(k_split_uv<true>, grid_of(w / 2, h / 2), dim3(256), 0, stream_of(stream), w / 2, h / 2, u, uStride, v, vStride,
                       uv, uvStride);
    MVX_LAUNCH_CHECK();
    return MVX_SUCCESS;
}

extern "C" int mvx_FormatConvert_UV12_IUV(mvx_stream_t stream, uint32_t cw, uint32_t ch, uint8_t *uv, uint32_t uvStride,
                                          const uint8_t *u, uint32_t uStride, const uint8_t *v, uint32_t vStride)
{
    if (!cw || !ch) return MVX_SUCCESS;
    if (!plane_ok(uv, uvStride, 2ull * cw) || !plane_ok(u, uStride, cw) || !plane_ok(v, vStride, cw)) return MVX_ERROR_INVALID_PARAMETERS;
    hipLaunchKernelGGL(k_merge_uv, grid_of(cw, ch), dim3(256), 0, stream_of(stream), cw, ch, uv, uvStride, u, uStride, v, vStride);
    MVX_LAUNCH_CHECK();
    return MVX_SUCCESS;
}

extern "C" int mvx_ScaleUp2x2_U8_U8(mvx_stream_t stream, uint32_t dw, uint32_t dh, uint8_t *dst, uint32_t dstStride,
                                    const uint8_t *src, uint32_t srcStride)
{
    if (!dw || !dh) return MVX_SUCCESS;
    if ((dw & 1u) || (dh & 1u) || !plane_ok(dst, dstStride, dw) || !plane_ok(src, srcStride, dw / 2)) return MVX_ERROR_INVALID_PARAMETERS;
    hipLaunchKernelGGL(k_scale_up2x2, grid_of(dw / 2, dh / 2), dim3(256), 0, stream_of(stream), dw, dh, dst, dstStride, src, srcStride);
    MVX_LAUNCH_CHECK();
    return MVX_SUCCESS;
}
