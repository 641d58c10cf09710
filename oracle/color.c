/* oracle/color.c -- CPU restatement of the reference's ColorConvert kernels outside the RGB -> IYUV /
 * IYUV, NV12 -> RGB pair (those are in geometric.c).  Test infrastructure only: the checker the GPU
 * kernels are compared with, never part of the product.
 *
 * Planes are passed with their row strides; 4:2:0 chroma planes are w/2 x h/2 samples (NV12 / NV21:
 * interleaved pairs), 4:2:2 images (UYVY, YUYV) hold 2 bytes per pixel.  Sizes as the reference
 * accepts them: 4:2:0 and 4:2:2 widths even, 4:2:0 heights even.
 *
 * Pinned by: bit-exact against oracle/_ref (every conversion the reference's CPU engine supports, on
 * random images whose widths cover the SSE bodies and the scalar tails of every kernel); closed forms
 * (grey chroma 128 gives R = G = B = Y; RGB -> RGBX -> RGB is the identity). */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "oracle.h"

static uint8_t clamp_trunc(float f) { return (uint8_t)fminf(fmaxf(f, 0.0f), 255.0f); }
static int sat(int v, int hi) { return v < 0 ? 0 : (v > hi ? hi : v); }
/* _mm_cvtps_epi32 under the default MXCSR mode: nearest, ties to even (all values here are small) */
static int rne(float f) { return (int)nearbyintf(f); }
/* (vx_uint8) of a double as gcc compiles it for x86: cvttsd2si to int, then the low byte */
static uint8_t u8_of(double d) { return (uint8_t)(int)d; }

/* ---------------------------------------------------------------- YUV -> RGB / RGBX --------------
 * HafCpu_ColorConvert_RGB_UYVY / _RGB_YUYV (ago_haf_cpu_color_convert.cpp:280-471), _RGBX_UYVY /
 * _RGBX_YUYV (:914-1273), _RGB_NV21 (:769-913), _RGBX_IYUV / _RGBX_NV12 / _RGBX_NV21 (:1274-1732).
 * Per pixel, u = U - 128 and v = V - 128 in float:
 *   R = Y + v*1.5748, B = Y + u*1.8556, and G = Y - (u*0.1873 + v*0.4681) -- the SSE bodies add the
 *   weight vectors (u*wU + v*wV, whose 0-weight terms change nothing) and the 4:2:0 scalar tails group
 *   the same way -- except in the 4:2:2 scalar tails (columns w & ~7 .. w-1), which compute
 *   G = (Y - u*0.1873) - v*0.4681;
 * each channel truncated and saturated to [0, 255] (the bodies truncate then saturate, the tails
 * clamp then truncate: the same bytes); RGBX adds X = 255. */
void oracle_cc_rgb_from_yuv(int fmt, int out_bpp, int w, int h, uint8_t *dst, int dst_stride, const uint8_t *p0, int s0,
                            const uint8_t *p1, int s1, const uint8_t *p2, int s2)
{
    const int body422 = w & ~7;
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            int Y, U, V;
            const long c = (long)(y / 2) * s1 + (x / 2);
            switch (fmt) {
            case ORACLE_FMT_IYUV: Y = p0[(long)y * s0 + x]; U = p1[c]; V = p2[(long)(y / 2) * s2 + (x / 2)]; break;
            case ORACLE_FMT_NV12: Y = p0[(long)y * s0 + x]; U = p1[(long)(y / 2) * s1 + 2 * (x / 2)]; V = p1[(long)(y / 2) * s1 + 2 * (x / 2) + 1]; break;
            case ORACLE_FMT_NV21: Y = p0[(long)y * s0 + x]; V = p1[(long)(y / 2) * s1 + 2 * (x / 2)]; U = p1[(long)(y / 2) * s1 + 2 * (x / 2) + 1]; break;
            case ORACLE_FMT_UYVY: {
                const uint8_t *g = p0 + (long)y * s0 + 4 * (x / 2);
                U = g[0]; Y = g[1 + 2 * (x & 1)]; V = g[2];
                break;
            }
            default: {   /* YUYV */
                const uint8_t *g = p0 + (long)y * s0 + 4 * (x / 2);
                Y = g[2 * (x & 1)]; U = g[1]; V = g[3];
                break;
            }
            }
            const float fy = (float)Y, u = (float)U - 128.0f, v = (float)V - 128.0f;
            const int tail422 = (fmt == ORACLE_FMT_UYVY || fmt == ORACLE_FMT_YUYV) && x >= body422;
            const float g = tail422 ? (fy - u * 0.1873f) - v * 0.4681f : fy - ((u * 0.1873f) + (v * 0.4681f));
            uint8_t *d = dst + (long)y * dst_stride + (long)out_bpp * x;
            d[0] = clamp_trunc(fy + v * 1.5748f);
            d[1] = clamp_trunc(g);
            d[2] = clamp_trunc(fy + u * 1.8556f);
            if (out_bpp == 4) d[3] = 255;
        }
}

/* ---------------------------------------------------------------- RGB / RGBX -> 4:2:0 -----------
 * HafCpu_ColorConvert_NV12_RGB (:2707-2912, the interleaved form of IYUV_RGB, :2492-2705, which
 * oracle_color_iyuv_from_rgb restates), HafCpu_ColorConvert_IYUV_RGBX / _NV12_RGBX (:3868-4281).
 * Per 2x2 block, columns below w & ~3 follow the SSE body, the last two columns when w % 4 == 2 the
 * scalar tail.
 *   RGB body:  Y = trunc(((R*.2126f) + G*.7152f) + B*.0722f); u = trunc(Uf + 128) per pixel with
 *              Uf = ((R*-.1146f) + G*-.3854f) + B*.5f (V: .5f, -.4542f, -.0458f); chroma =
 *              ((u00+u10+1)>>1 + (u01+u11+1)>>1 + 1) >> 1, saturated;
 *   RGBX body: Y = rne(Yf); u = rne(Uf) + 128; chroma = ((u00+u10+1)>>1 + (u01+u11+1)>>1) >> 1 (no
 *              final rounding), saturated;
 *   tails:     RGB  Y = (u8)((double)((R*.2126f) + G*.7152f) + B*0.0722), chroma = (u8)(float sum of the
 *              four pixels' Uf + 128, row-major, / 4);
 *              RGBX the same with round() before each cast. */
void oracle_cc_yuv420_from_rgb(int in_bpp, int nv12, int w, int h, const uint8_t *src, int src_stride, uint8_t *yp, int ys,
                               uint8_t *up, int us, uint8_t *vp, int vs)
{
    const int aligned = w & ~3;
    for (int r = 0; r < h; r += 2)
        for (int x = 0; x < w; x += 2) {
            float R[4], G[4], B[4];
            for (int k = 0; k < 4; k++) {
                const uint8_t *p = src + (long)(r + (k >> 1)) * src_stride + (long)in_bpp * (x + (k & 1));
                R[k] = p[0]; G[k] = p[1]; B[k] = p[2];
            }
            int uo, vo;
            if (x < aligned) {
                int uu[4], vv[4];
                for (int k = 0; k < 4; k++) {
                    const float yf = ((R[k] * 0.2126f) + (G[k] * 0.7152f)) + (B[k] * 0.0722f);
                    const float uf = ((R[k] * -0.1146f) + (G[k] * -0.3854f)) + (B[k] * 0.5f);
                    const float vf = ((R[k] * 0.5f) + (G[k] * -0.4542f)) + (B[k] * -0.0458f);
                    int yv;
                    if (in_bpp == 3) {
                        yv = sat((int)yf, 255);
                        uu[k] = sat((int)(uf + 128.0f), 65535);
                        vv[k] = sat((int)(vf + 128.0f), 65535);
                    } else {
                        yv = sat(rne(yf), 255);
                        uu[k] = sat(rne(uf) + 128, 65535);
                        vv[k] = sat(rne(vf) + 128, 65535);
                    }
                    yp[(long)(r + (k >> 1)) * ys + x + (k & 1)] = (uint8_t)yv;
                }
                const int round = in_bpp == 3 ? 1 : 0;
                uo = sat((((uu[0] + uu[2] + 1) >> 1) + ((uu[1] + uu[3] + 1) >> 1) + round) >> 1, 255);
                vo = sat((((vv[0] + vv[2] + 1) >> 1) + ((vv[1] + vv[3] + 1) >> 1) + round) >> 1, 255);
            } else {
                float U = 0.0f, V = 0.0f;
                for (int k = 0; k < 4; k++) {
                    const double yd = (double)((R[k] * 0.2126f) + (G[k] * 0.7152f)) + (double)B[k] * 0.0722;
                    yp[(long)(r + (k >> 1)) * ys + x + (k & 1)] = in_bpp == 3 ? u8_of(yd) : u8_of(round(yd));
                    const float uf = ((R[k] * -0.1146f) + (G[k] * -0.3854f)) + (B[k] * 0.5f) + 128.0f;
                    const float vf = ((R[k] * 0.5f) + (G[k] * -0.4542f)) + (B[k] * -0.0458f) + 128.0f;
                    U = k ? U + uf : uf;
                    V = k ? V + vf : vf;
                }
                U /= 4.0f;
                V /= 4.0f;
                uo = in_bpp == 3 ? u8_of(U) : u8_of(round(U));
                vo = in_bpp == 3 ? u8_of(V) : u8_of(round(V));
            }
            if (nv12) {
                uint8_t *c = up + (long)(r >> 1) * us + x;
                c[0] = (uint8_t)uo;
                c[1] = (uint8_t)vo;
            } else {
                up[(long)(r >> 1) * us + (x >> 1)] = (uint8_t)uo;
                vp[(long)(r >> 1) * vs + (x >> 1)] = (uint8_t)vo;
            }
        }
}

/* ---------------------------------------------------------------- RGB / RGBX -> YUV4 -------------
 * HafCpu_ColorConvert_YUV4_RGB (:4282-4528) and _YUV4_RGBX (:3740-3867).  Columns below w & ~15 (SSE
 * body): Y = rne(((R*.2126f) + G*.7152f) + B*.0722f), U = rne(((R*-.1146f) + G*-.3854f) + B*.5f) + 128,
 * V = rne(((R*.5f) + G*-.4542f) + B*-.0458f) + 128, each saturated to [0, 255].  Scalar tail:
 *   Y = (u8)round((double)((R*.2126f) + G*.7152f) + B*0.0722)
 *   U = (u8)round(((double)(R*-.1146f) + G*-0.3854) + (double)(B*.5f) + 128)     (G's weight is double)
 *   V = (u8)round((float)(((R*.5f) + G*-.4542f) + B*-.0458f) + 128.f)
 * with (u8) the low byte of the integer conversion (256 becomes 0). */
void oracle_cc_yuv4_from_rgb(int in_bpp, int w, int h, const uint8_t *src, int src_stride, uint8_t *yp, int ys, uint8_t *up,
                             int us, uint8_t *vp, int vs)
{
    const int aligned = w & ~15;
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            const uint8_t *p = src + (long)y * src_stride + (long)in_bpp * x;
            const float R = p[0], G = p[1], B = p[2];
            uint8_t Y, U, V;
            if (x < aligned) {
                Y = (uint8_t)sat(rne(((R * 0.2126f) + (G * 0.7152f)) + (B * 0.0722f)), 255);
                U = (uint8_t)sat(rne(((R * -0.1146f) + (G * -0.3854f)) + (B * 0.5f)) + 128, 255);
                V = (uint8_t)sat(rne(((R * 0.5f) + (G * -0.4542f)) + (B * -0.0458f)) + 128, 255);
            } else {
                Y = u8_of(round((double)((R * 0.2126f) + (G * 0.7152f)) + (double)B * 0.0722));
                U = u8_of(round((((double)(R * -0.1146f) + (double)G * -0.3854) + (double)(B * 0.5f)) + 128.0));
                V = u8_of(round((double)((((R * 0.5f) + (G * -0.4542f)) + (B * -0.0458f)) + 128.0f)));
            }
            yp[(long)y * ys + x] = Y;
            up[(long)y * us + x] = U;
            vp[(long)y * vs + x] = V;
        }
}

/* ---------------------------------------------------------------- RGB <-> RGBX -------------------
 * HafCpu_ColorConvert_RGB_RGBX (:2330-2398, drops X) and _RGBX_RGB (:2399-2491, X = 255). */
void oracle_cc_rgb_rgbx(int to_rgbx, int w, int h, const uint8_t *src, int src_stride, uint8_t *dst, int dst_stride)
{
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            const uint8_t *s = src + (long)y * src_stride + (long)(to_rgbx ? 3 : 4) * x;
            uint8_t *d = dst + (long)y * dst_stride + (long)(to_rgbx ? 4 : 3) * x;
            d[0] = s[0]; d[1] = s[1]; d[2] = s[2];
            if (to_rgbx) d[3] = 255;
        }
}

/* ---------------------------------------------------------------- YUV -> YUV --------------------
 * The byte-moving conversions: HafCpu_FormatConvert_IYUV_UYVY / _IYUV_YUYV / _NV12_UYVY / _NV12_YUYV
 * (:55-279, :1733-2329: luma copied, each chroma sample the floor average (a + b) >> 1 of the two rows'
 * samples -- the SSE bodies' rounding average corrected down by the low bit of a ^ b), and the
 * divisions into a luma ChannelCopy plus HafCpu_FormatConvert_UV12_IUV (IYUV -> NV12, :4663-4732:
 * interleave), _IUV_UV12 (NV12 / NV21 -> IYUV, :4529-4662: de-interleave), _UV_UV12 (NV12 / NV21 ->
 * YUV4, :4733-: each sample to a 2x2 block) and HafCpu_ScaleUp2x2_U8_U8 (IYUV -> YUV4,
 * ago_haf_cpu_geometric.cpp:2420-: the same 2x2 replication), ago_drama_divide.cpp:157-325.
 * `in` / `out` hold every plane tightly packed, one after the other.  Returns -1 for a pair the
 * reference does not convert. */
int oracle_cc_format(uint32_t in_fmt, uint32_t out_fmt, int w, int h, const uint8_t *in, uint8_t *out)
{
#define FOURCC(a, b, c, d) ((uint32_t)(a) | (uint32_t)(b) << 8 | (uint32_t)(c) << 16 | (uint32_t)(d) << 24)
    const uint32_t IYUV = FOURCC('I', 'Y', 'U', 'V'), NV12 = FOURCC('N', 'V', '1', '2'), NV21 = FOURCC('N', 'V', '2', '1');
    const uint32_t YUV4 = FOURCC('Y', 'U', 'V', '4'), UYVY = FOURCC('U', 'Y', 'V', 'Y'), YUYV = FOURCC('Y', 'U', 'Y', 'V');
#undef FOURCC
    const int cw = w / 2, ch = h / 2;
    uint8_t *oy = out;
    if (in_fmt == UYVY || in_fmt == YUYV) {
        if (out_fmt != IYUV && out_fmt != NV12) return -1;
        const int yo = in_fmt == UYVY ? 1 : 0, uo = in_fmt == UYVY ? 0 : 1, vo = in_fmt == UYVY ? 2 : 3;
        for (int y = 0; y < h; y++)
            for (int x = 0; x < w; x++) oy[(long)y * w + x] = in[(long)y * 2 * w + 4 * (x / 2) + yo + 2 * (x & 1)];
        for (int j = 0; j < ch; j++)
            for (int i = 0; i < cw; i++) {
                const uint8_t *g0 = in + (long)(2 * j) * 2 * w + 4 * i, *g1 = g0 + 2 * w;
                const uint8_t u = (uint8_t)((g0[uo] + g1[uo]) >> 1), v = (uint8_t)((g0[vo] + g1[vo]) >> 1);
                if (out_fmt == IYUV) {
                    out[(long)w * h + (long)j * cw + i] = u;
                    out[(long)w * h + (long)cw * ch + (long)j * cw + i] = v;
                } else {
                    out[(long)w * h + (long)j * w + 2 * i] = u;
                    out[(long)w * h + (long)j * w + 2 * i + 1] = v;
                }
            }
        return 0;
    }
    if (in_fmt != IYUV && in_fmt != NV12 && in_fmt != NV21) return -1;
    if (out_fmt == in_fmt || !(out_fmt == IYUV || out_fmt == NV12 || out_fmt == YUV4) || (in_fmt == NV21 && out_fmt == NV12))
        return -1;
    memcpy(oy, in, (size_t)w * h);
    const uint8_t *ci = in + (long)w * h;
    for (int j = 0; j < ch; j++)
        for (int i = 0; i < cw; i++) {
            uint8_t u, v;
            if (in_fmt == IYUV) {
                u = ci[(long)j * cw + i];
                v = ci[(long)cw * ch + (long)j * cw + i];
            } else {
                const uint8_t a = ci[(long)j * w + 2 * i], b = ci[(long)j * w + 2 * i + 1];
                u = in_fmt == NV12 ? a : b;
                v = in_fmt == NV12 ? b : a;
            }
            uint8_t *co = out + (long)w * h;
            if (out_fmt == IYUV) {
                co[(long)j * cw + i] = u;
                co[(long)cw * ch + (long)j * cw + i] = v;
            } else if (out_fmt == NV12) {
                co[(long)j * w + 2 * i] = u;
                co[(long)j * w + 2 * i + 1] = v;
            } else {
                for (int k = 0; k < 4; k++) {
                    const long o = (long)(2 * j + (k >> 1)) * w + 2 * i + (k & 1);
                    co[o] = u;
                    co[(long)w * h + o] = v;
                }
            }
        }
    return 0;
}

/* ColorConvert_Y_RGBX (ago_haf_cpu_color_convert.cpp:3367-3487), the luma kernel the reference keeps when
 * only the Y channel of an RGBX conversion is read: columns below w & ~15 truncate the float sum
 * ((R*.2126f) + G*.7152f) + B*.0722f (SSE body); the rest truncate (double)((R*.2126f) + G*.7152f) +
 * B*0.0722 (scalar tail, a double constant). */
void oracle_cc_y_rgbx(int w, int h, const uint8_t *src, int src_stride, uint8_t *dst, int dst_stride)
{
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            const uint8_t *p = src + (long)y * src_stride + 4L * x;
            const float R = p[0], G = p[1], B = p[2];
            dst[(long)y * dst_stride + x] = x < (w & ~15) ? (uint8_t)sat((int)(((R * 0.2126f) + (G * 0.7152f)) + (B * 0.0722f)), 255)
                                                          : u8_of((double)((R * 0.2126f) + (G * 0.7152f)) + (double)B * 0.0722);
        }
}
