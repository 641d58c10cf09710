// vx_kernels.cpp -- the engine's kernel table: Khronos vision nodes and the AMD low-level kernels
// that GDF files may name directly, each with its validate / initialize / execute commands.
//
// Table rows follow ago_kernel_list.cpp:160-520 (names "org.khronos.openvx.<name>" and
// "com.amd.openvx.<Name>", kernel enums, argument directions and order); the command protocol is
// that of the agoKernel_* / ovxKernel_* functions in ago_kernel_api.cpp (validate :ago_kernel_cmd_
// validate, initialize :ago_kernel_cmd_initialize, hip_execute :ago_kernel_cmd_hip_execute).  The
// Khronos nodes are split into AMD kernels the way the reference's graph divider does it
// (ago_drama_divide.cpp, cited per node below), but inside one execute command: every kernel is a
// gfx950 launch behind mvx_kernels.h on the graph's stream, including the Canny trace and the
// Harris sort-and-pick that the reference runs on the CPU.
#include <algorithm>
#include <cmath>

#include "engine.hpp"

using namespace mvxe;

namespace {

constexpr vx_enum IN = VX_INPUT, OUT = VX_OUTPUT;


_vx_image *img(_vx_node *n, int i) { return image_of(n->params[i]); }
_vx_scalar *scl(_vx_node *n, int i) { return is_ref(n->params[i], VX_TYPE_SCALAR) ? (_vx_scalar *)n->params[i] : nullptr; }
uint8_t *P(_vx_image *im, uint32_t plane = 0) { return im->plane_ptr(plane); }
uint32_t S(_vx_image *im, uint32_t plane = 0) { return im->planes[plane].stride; }

vx_status fail(_vx_node *n, vx_status s, const char *what)
{
    log(n, s, "ERROR: %s: %s\n", n->kernel->name, what);
    return s;
}

// an output image: virtual images with unset size/format take the given meta data; others must
// already have it (ago_kernel_api.cpp ValidateArguments_* helpers)
vx_status set_output(_vx_node *n, int i, uint32_t w, uint32_t h, vx_df_image fmt)
{
    _vx_image *o = img(n, i);
    if (!o) return fail(n, VX_ERROR_INVALID_PARAMETERS, "output is not an image");
    if (o->isVirtual && (o->width == 0 || o->height == 0 || o->format == VX_DF_IMAGE_VIRT)) {
        if (o->width == 0) o->width = w;
        if (o->height == 0) o->height = h;
        if (o->format == VX_DF_IMAGE_VIRT) o->format = fmt;
        if (!o->layout()) return fail(n, VX_ERROR_INVALID_FORMAT, "bad output format");
    }
    if (o->format != fmt) return fail(n, VX_ERROR_INVALID_FORMAT, "output format");
    if (o->width != w || o->height != h) return fail(n, VX_ERROR_INVALID_DIMENSION, "output dimensions");
    return VX_SUCCESS;
}

vx_status need_input(_vx_node *n, int i, vx_df_image fmt)
{
    _vx_image *in = img(n, i);
    if (!in) return fail(n, VX_ERROR_INVALID_PARAMETERS, "input is not an image");
    if (in->format != fmt) return fail(n, VX_ERROR_INVALID_FORMAT, "input format");
    if (in->width == 0 || in->height == 0) return fail(n, VX_ERROR_INVALID_DIMENSION, "input dimensions");
    return VX_SUCCESS;
}

vx_status border_undefined(_vx_node *n)
{
    if (n->border.mode != VX_BORDER_UNDEFINED)
        return fail(n, VX_ERROR_NOT_SUPPORTED, "only the UNDEFINED border mode is provided for this node");
    return VX_SUCCESS;
}

vx_status rc(_vx_node *n, int r)
{
    if (r == 0) return VX_SUCCESS;
    log(n, r, "ERROR: %s: kernel launch returned %d\n", n->kernel->name, r);
    return r;
}

// ---------------------------------------------------------------------- 3x3 filters (A4, A5) ----
// gaussian_3x3 / box_3x3 / median_3x3 -> one AMD kernel each (ago_drama_divide.cpp:835-875)
typedef int (*Filter3x3)(mvx_stream_t, uint32_t, uint32_t, uint8_t *, uint32_t, const uint8_t *, uint32_t);

template <int IO>  // IO 0: Khronos order (in, out); 1: AMD order (out, in)
vx_status v_filter(_vx_node *n)
{
    const int i = IO ? 1 : 0, o = IO ? 0 : 1;
    vx_status s = need_input(n, i, VX_DF_IMAGE_U8);
    if (!s) s = border_undefined(n);
    if (!s) s = set_output(n, o, img(n, i)->width, img(n, i)->height, VX_DF_IMAGE_U8);
    return s;
}

template <int IO, Filter3x3 F> vx_status x_filter(_vx_node *n, hipStream_t st)
{
    _vx_image *in = img(n, IO ? 1 : 0), *out = img(n, IO ? 0 : 1);
    return rc(n, F(st, out->width, out->height, P(out), S(out), P(in), S(in)));
}

// ---------------------------------------------------------------------- Sobel (A6) ---------------
// sobel_3x3 -> SOBEL_S16S16_U8_3x3_GXY, or _GX / _GY with one output (ago_drama_divide.cpp:509-531)
vx_status v_sobel(_vx_node *n)
{
    vx_status s = need_input(n, 0, VX_DF_IMAGE_U8);
    if (s) return s;
    if (!n->params[1] && !n->params[2]) return fail(n, VX_ERROR_INVALID_PARAMETERS, "no output");
    if ((s = border_undefined(n))) return s;
    for (int i = 1; i <= 2; i++)
        if (n->params[i] && (s = set_output(n, i, img(n, 0)->width, img(n, 0)->height, VX_DF_IMAGE_S16))) return s;
    return VX_SUCCESS;
}

// the output image of an absorbed Magnitude / Phase node (Khronos order: out last; AMD order: out first)
_vx_image *gradpair_out(_vx_node *m)
{
    return img(m, (m->kernel->id == AMD_MAGNITUDE_S16_S16S16 || m->kernel->id == AMD_PHASE_U8_S16S16) ? 0 : 2);
}

// a Sobel node whose virtual gradients only a Magnitude and/or a Phase node read runs the merged kernel
// the reference's node-merge rules choose (ago_drama_remove.cpp:155-182) and writes their outputs
vx_status x_sobel_merged(_vx_node *n, hipStream_t st, _vx_image *in)
{
    _vx_image *m = n->mergedMag ? gradpair_out(n->mergedMag) : nullptr, *p = n->mergedPhase ? gradpair_out(n->mergedPhase) : nullptr;
    const uint32_t w = in->width, h = in->height;
    vx_status s;
    if (m && p) s = rc(n, mvx_SobelMagnitudePhase_S16U8_U8_3x3(st, w, h, (int16_t *)P(m), S(m), P(p), S(p), P(in), S(in)));
    else if (m) s = rc(n, mvx_SobelMagnitude_S16_U8_3x3(st, w, h, (int16_t *)P(m), S(m), P(in), S(in)));
    else s = rc(n, mvx_SobelPhase_U8_U8_3x3(st, w, h, P(p), S(p), P(in), S(in)));
    for (_vx_node *a : {n->mergedMag, n->mergedPhase}) if (a) a->status = s;
    return s;
}

vx_status x_sobel(_vx_node *n, hipStream_t st)
{
    _vx_image *in = img(n, 0), *gx = img(n, 1), *gy = img(n, 2);
    if (n->mergedMag || n->mergedPhase) return x_sobel_merged(n, st, in);
    return rc(n, mvx_Sobel_S16S16_U8_3x3_GXY(st, in->width, in->height, gx ? (int16_t *)P(gx) : nullptr, gx ? S(gx) : 0,
                                             gy ? (int16_t *)P(gy) : nullptr, gy ? S(gy) : 0, P(in), S(in)));
}

vx_status v_amd_sobel(_vx_node *n)
{
    vx_status s = need_input(n, 2, VX_DF_IMAGE_U8);
    if (!s) s = set_output(n, 0, img(n, 2)->width, img(n, 2)->height, VX_DF_IMAGE_S16);
    if (!s) s = set_output(n, 1, img(n, 2)->width, img(n, 2)->height, VX_DF_IMAGE_S16);
    return s;
}

vx_status x_amd_sobel(_vx_node *n, hipStream_t st)
{
    _vx_image *gx = img(n, 0), *gy = img(n, 1), *in = img(n, 2);
    if (n->mergedMag || n->mergedPhase) return x_sobel_merged(n, st, in);
    return rc(n, mvx_Sobel_S16S16_U8_3x3_GXY(st, in->width, in->height, (int16_t *)P(gx), S(gx), (int16_t *)P(gy), S(gy), P(in), S(in)));
}

// AMD merged forms named directly (ago_kernel_list.cpp:416-418; wrappers ago_kernel_api.cpp:15958-16110):
// SobelMagnitude_S16_U8_3x3 (mag, in), SobelPhase_U8_U8_3x3 (phase, in),
// SobelMagnitudePhase_S16U8_U8_3x3 (mag, phase, in)
template <int OUTS> vx_status v_amd_sobel_merged(_vx_node *n)
{
    vx_status s = need_input(n, OUTS, VX_DF_IMAGE_U8);
    if (s) return s;
    _vx_image *in = img(n, OUTS);
    if (n->kernel->id == AMD_SOBEL_PHASE_U8_U8_3x3) return set_output(n, 0, in->width, in->height, VX_DF_IMAGE_U8);
    s = set_output(n, 0, in->width, in->height, VX_DF_IMAGE_S16);
    if (!s && OUTS == 2) s = set_output(n, 1, in->width, in->height, VX_DF_IMAGE_U8);
    return s;
}

template <vx_enum ID> vx_status x_amd_sobel_merged(_vx_node *n, hipStream_t st)
{
    if (ID == AMD_SOBEL_MAGNITUDE_PHASE_S16U8_U8_3x3) {
        _vx_image *m = img(n, 0), *p = img(n, 1), *in = img(n, 2);
        return rc(n, mvx_SobelMagnitudePhase_S16U8_U8_3x3(st, in->width, in->height, (int16_t *)P(m), S(m), P(p), S(p), P(in), S(in)));
    }
    _vx_image *o = img(n, 0), *in = img(n, 1);
    if (ID == AMD_SOBEL_MAGNITUDE_S16_U8_3x3)
        return rc(n, mvx_SobelMagnitude_S16_U8_3x3(st, in->width, in->height, (int16_t *)P(o), S(o), P(in), S(in)));
    return rc(n, mvx_SobelPhase_U8_U8_3x3(st, in->width, in->height, P(o), S(o), P(in), S(in)));
}

// ---------------------------------------------------------------------- Magnitude / Phase (A7, A8)
// magnitude -> MAGNITUDE_S16_S16S16, phase -> PHASE_U8_S16S16, output first (ago_drama_divide.cpp:533-561)
template <int IO, vx_df_image OF> vx_status v_gradpair(_vx_node *n)
{
    const int a = IO ? 1 : 0, b = IO ? 2 : 1, o = IO ? 0 : 2;
    vx_status s = need_input(n, a, VX_DF_IMAGE_S16);
    if (!s) s = need_input(n, b, VX_DF_IMAGE_S16);
    if (s) return s;
    if (img(n, a)->width != img(n, b)->width || img(n, a)->height != img(n, b)->height)
        return fail(n, VX_ERROR_INVALID_DIMENSION, "gradient sizes differ");
    return set_output(n, o, img(n, a)->width, img(n, a)->height, OF);
}

template <int IO> vx_status x_magnitude(_vx_node *n, hipStream_t st)
{
    _vx_image *gx = img(n, IO ? 1 : 0), *gy = img(n, IO ? 2 : 1), *m = img(n, IO ? 0 : 2);
    return rc(n, mvx_Magnitude_S16_S16S16(st, m->width, m->height, (int16_t *)P(m), S(m), (const int16_t *)P(gx), S(gx),
                                          (const int16_t *)P(gy), S(gy)));
}

template <int IO> vx_status x_phase(_vx_node *n, hipStream_t st)
{
    _vx_image *gx = img(n, IO ? 1 : 0), *gy = img(n, IO ? 2 : 1), *p = img(n, IO ? 0 : 2);
    return rc(n, mvx_Phase_U8_S16S16(st, p->width, p->height, P(p), S(p), (const int16_t *)P(gx), S(gx), (const int16_t *)P(gy), S(gy)));
}

// ---------------------------------------------------------------------- ColorConvert (A2) --------
// The conversions the reference's divider accepts (ago_drama_divide.cpp:40-325) and the kernels its
// node-merge rules leave (ago_drama_remove.cpp:38-100):
//   YUV (UYVY, YUYV, IYUV, NV12, NV21) -> RGB / RGBX: COLOR_CONVERT_RGB[X]_<fmt>;
//   RGB / RGBX -> IYUV, NV12, YUV4: the merged COLOR_CONVERT_<fmt>_RGB[X] kernel when every plane is used,
//     only the luma kernel (COLOR_CONVERT_Y_RGB[X]) when the output is virtual and read only through its
//     Y channel (the divider's chroma kernels have no readers and are removed, ago_drama_remove.cpp:989);
//   RGB <-> RGBX; UYVY / YUYV -> IYUV / NV12 (FORMAT_CONVERT_*); IYUV -> NV12 (a luma copy plus
//     UV12_IUV), IYUV -> YUV4 (copy plus ScaleUp2x2 of U and V), NV12 / NV21 -> IYUV (copy plus IUV_UV12
//     with the chroma outputs swapped for NV21) and -> YUV4 (copy plus UV_UV12).
// 4:2:2 images need an even width, 4:2:0 images an even width and height.
// true when every node of the graph that reads `yuv` is a channel_extract of its Y channel: the RGB
// conversion then needs only its luma kernel
bool luma_only_readers(_vx_node *n, _vx_image *yuv)
{
    for (_vx_node *m : n->graph->nodes)
        for (size_t i = 0; i < m->params.size(); i++) {
            if (m->params[i] != yuv || m->kernel->params[i].dir == VX_OUTPUT) continue;
            if (m->kernel->id != VX_KERNEL_CHANNEL_EXTRACT) return false;
            _vx_scalar *ch = scl(m, 1);
            if (!ch || (ch->value.e != VX_CHANNEL_Y && ch->value.e != VX_CHANNEL_0)) return false;
        }
    return true;
}

enum : uint32_t { COLOR_ALL_PLANES = 0, COLOR_LUMA_ONLY = 1 };

bool is_rgb(vx_df_image f) { return f == VX_DF_IMAGE_RGB || f == VX_DF_IMAGE_RGBX; }
bool is_422(vx_df_image f) { return f == VX_DF_IMAGE_UYVY || f == VX_DF_IMAGE_YUYV; }
bool is_420(vx_df_image f) { return f == VX_DF_IMAGE_IYUV || f == VX_DF_IMAGE_NV12 || f == VX_DF_IMAGE_NV21; }

bool color_pair_ok(vx_df_image in, vx_df_image out)
{
    if (is_rgb(in)) return (is_rgb(out) && out != in) || out == VX_DF_IMAGE_IYUV || out == VX_DF_IMAGE_NV12 || out == VX_DF_IMAGE_YUV4;
    if (is_422(in)) return is_rgb(out) || out == VX_DF_IMAGE_IYUV || out == VX_DF_IMAGE_NV12;
    if (in == VX_DF_IMAGE_IYUV) return is_rgb(out) || out == VX_DF_IMAGE_NV12 || out == VX_DF_IMAGE_YUV4;
    if (in == VX_DF_IMAGE_NV12 || in == VX_DF_IMAGE_NV21) return is_rgb(out) || out == VX_DF_IMAGE_IYUV || out == VX_DF_IMAGE_YUV4;
    return false;
}

vx_status v_color(_vx_node *n)
{
    _vx_image *in = img(n, 0), *o = img(n, 1);
    if (!in) return fail(n, VX_ERROR_INVALID_PARAMETERS, "input is not an image");
    if (!o) return fail(n, VX_ERROR_INVALID_PARAMETERS, "output is not an image");
    if (in->width == 0 || in->height == 0) return fail(n, VX_ERROR_INVALID_DIMENSION, "input dimensions");
    const vx_df_image want = o->format != VX_DF_IMAGE_VIRT ? o->format : is_rgb(in->format) ? VX_DF_IMAGE_IYUV : VX_DF_IMAGE_RGB;
    if (!color_pair_ok(in->format, want)) return fail(n, VX_ERROR_NOT_SUPPORTED, "color conversion between these formats");
    const bool yuv420 = is_420(in->format) || is_420(want), yuv422 = is_422(in->format);
    vx_status s = set_output(n, 1, in->width, in->height, want);
    if (s) return s;
    n->variant = (is_rgb(in->format) && !is_rgb(want) && o->isVirtual && luma_only_readers(n, o)) ? COLOR_LUMA_ONLY : COLOR_ALL_PLANES;
    if (n->variant == COLOR_LUMA_ONLY) return VX_SUCCESS;   // the luma kernels take any size
    if ((yuv420 || yuv422) && (in->width & 1u)) return fail(n, VX_ERROR_INVALID_DIMENSION, "YUV widths must be even");
    if (yuv420 && (in->height & 1u)) return fail(n, VX_ERROR_INVALID_DIMENSION, "YUV 4:2:0 heights must be even");
    return VX_SUCCESS;
}

vx_status x_color(_vx_node *n, hipStream_t st)
{
    _vx_image *in = img(n, 0), *o = img(n, 1);
    const uint32_t w = o->width, h = o->height;
    const vx_df_image fi = in->format, fo = o->format;
    if (n->variant == COLOR_LUMA_ONLY)
        return rc(n, fi == VX_DF_IMAGE_RGB ? mvx_ColorConvert_Y_RGB(st, w, h, P(o), S(o), P(in), S(in))
                                           : mvx_ColorConvert_Y_RGBX(st, w, h, P(o), S(o), P(in), S(in)));
    // YUV -> RGB / RGBX
    if (is_rgb(fo) && !is_rgb(fi)) {
        const bool x = fo == VX_DF_IMAGE_RGBX;
        switch (fi) {
        case VX_DF_IMAGE_UYVY:
            return rc(n, x ? mvx_ColorConvert_RGBX_UYVY(st, w, h, P(o), S(o), P(in), S(in)) : mvx_ColorConvert_RGB_UYVY(st, w, h, P(o), S(o), P(in), S(in)));
        case VX_DF_IMAGE_YUYV:
            return rc(n, x ? mvx_ColorConvert_RGBX_YUYV(st, w, h, P(o), S(o), P(in), S(in)) : mvx_ColorConvert_RGB_YUYV(st, w, h, P(o), S(o), P(in), S(in)));
        case VX_DF_IMAGE_IYUV:
            return rc(n, x ? mvx_ColorConvert_RGBX_IYUV(st, w, h, P(o), S(o), P(in, 0), S(in, 0), P(in, 1), S(in, 1), P(in, 2), S(in, 2))
                           : mvx_ColorConvert_RGB_IYUV(st, w, h, P(o), S(o), P(in, 0), S(in, 0), P(in, 1), S(in, 1), P(in, 2), S(in, 2)));
        case VX_DF_IMAGE_NV12:
            return rc(n, x ? mvx_ColorConvert_RGBX_NV12(st, w, h, P(o), S(o), P(in, 0), S(in, 0), P(in, 1), S(in, 1))
                           : mvx_ColorConvert_RGB_NV12(st, w, h, P(o), S(o), P(in, 0), S(in, 0), P(in, 1), S(in, 1)));
        default:
            return rc(n, x ? mvx_ColorConvert_RGBX_NV21(st, w, h, P(o), S(o), P(in, 0), S(in, 0), P(in, 1), S(in, 1))
                           : mvx_ColorConvert_RGB_NV21(st, w, h, P(o), S(o), P(in, 0), S(in, 0), P(in, 1), S(in, 1)));
        }
    }
    if (is_rgb(fi) && is_rgb(fo))
        return rc(n, fo == VX_DF_IMAGE_RGBX ? mvx_ColorConvert_RGBX_RGB(st, w, h, P(o), S(o), P(in), S(in))
                                            : mvx_ColorConvert_RGB_RGBX(st, w, h, P(o), S(o), P(in), S(in)));
    // RGB / RGBX -> YUV (every plane)
    if (is_rgb(fi)) {
        const bool x = fi == VX_DF_IMAGE_RGBX;
        if (fo == VX_DF_IMAGE_IYUV)
            return rc(n, x ? mvx_ColorConvert_IYUV_RGBX(st, w, h, P(o, 0), S(o, 0), P(o, 1), S(o, 1), P(o, 2), S(o, 2), P(in), S(in))
                           : mvx_ColorConvert_IYUV_RGB(st, w, h, P(o, 0), S(o, 0), P(o, 1), S(o, 1), P(o, 2), S(o, 2), P(in), S(in)));
        if (fo == VX_DF_IMAGE_NV12)
            return rc(n, x ? mvx_ColorConvert_NV12_RGBX(st, w, h, P(o, 0), S(o, 0), P(o, 1), S(o, 1), P(in), S(in))
                           : mvx_ColorConvert_NV12_RGB(st, w, h, P(o, 0), S(o, 0), P(o, 1), S(o, 1), P(in), S(in)));
        return rc(n, x ? mvx_ColorConvert_YUV4_RGBX(st, w, h, P(o, 0), S(o, 0), P(o, 1), S(o, 1), P(o, 2), S(o, 2), P(in), S(in))
                       : mvx_ColorConvert_YUV4_RGB(st, w, h, P(o, 0), S(o, 0), P(o, 1), S(o, 1), P(o, 2), S(o, 2), P(in), S(in)));
    }
    // 4:2:2 -> 4:2:0
    if (is_422(fi)) {
        const bool u = fi == VX_DF_IMAGE_UYVY;
        if (fo == VX_DF_IMAGE_IYUV)
            return rc(n, u ? mvx_FormatConvert_IYUV_UYVY(st, w, h, P(o, 0), S(o, 0), P(o, 1), S(o, 1), P(o, 2), S(o, 2), P(in), S(in))
                           : mvx_FormatConvert_IYUV_YUYV(st, w, h, P(o, 0), S(o, 0), P(o, 1), S(o, 1), P(o, 2), S(o, 2), P(in), S(in)));
        return rc(n, u ? mvx_FormatConvert_NV12_UYVY(st, w, h, P(o, 0), S(o, 0), P(o, 1), S(o, 1), P(in), S(in))
                       : mvx_FormatConvert_NV12_YUYV(st, w, h, P(o, 0), S(o, 0), P(o, 1), S(o, 1), P(in), S(in)));
    }
    // 4:2:0 -> 4:2:0 / YUV4: the luma plane copied, the chroma moved
    int r = mvx_ChannelCopy_U8_U8(st, w, h, P(o, 0), S(o, 0), P(in, 0), S(in, 0));
    if (r) return rc(n, r);
    if (fi == VX_DF_IMAGE_IYUV) {
        if (fo == VX_DF_IMAGE_NV12) return rc(n, mvx_FormatConvert_UV12_IUV(st, w / 2, h / 2, P(o, 1), S(o, 1), P(in, 1), S(in, 1), P(in, 2), S(in, 2)));
        r = mvx_ScaleUp2x2_U8_U8(st, w, h, P(o, 1), S(o, 1), P(in, 1), S(in, 1));
        return rc(n, r ? r : mvx_ScaleUp2x2_U8_U8(st, w, h, P(o, 2), S(o, 2), P(in, 2), S(in, 2)));
    }
    const bool nv21 = fi == VX_DF_IMAGE_NV21;   // NV21 holds V before U: the chroma outputs swap
    const int ou = nv21 ? 2 : 1, ov = nv21 ? 1 : 2;
    if (fo == VX_DF_IMAGE_IYUV)
        return rc(n, mvx_FormatConvert_IUV_UV12(st, w / 2, h / 2, P(o, ou), S(o, ou), P(o, ov), S(o, ov), P(in, 1), S(in, 1)));
    return rc(n, mvx_FormatConvert_UV_UV12(st, w, h, P(o, ou), S(o, ou), P(o, ov), S(o, ov), P(in, 1), S(in, 1)));
}

vx_status v_amd_color_y(_vx_node *n)
{
    vx_status s = need_input(n, 1, VX_DF_IMAGE_RGB);
    if (!s) s = set_output(n, 0, img(n, 1)->width, img(n, 1)->height, VX_DF_IMAGE_U8);
    return s;
}

vx_status x_amd_color_y(_vx_node *n, hipStream_t st)
{
    _vx_image *o = img(n, 0), *in = img(n, 1);
    return rc(n, mvx_ColorConvert_Y_RGB(st, o->width, o->height, P(o), S(o), P(in), S(in)));
}

// ---------------------------------------------------------------------- ChannelExtract (A3) ------
// channel_extract(Y) of a YUV image, and U / V of IYUV and YUV4 -> CHANNEL_COPY_U8_U8 of that plane
// (ago_drama_divide.cpp:342-414)
vx_status v_extract(_vx_node *n)
{
    _vx_image *in = img(n, 0);
    _vx_scalar *ch = scl(n, 1);
    if (!in || !ch) return fail(n, VX_ERROR_INVALID_PARAMETERS, "arguments");
    const vx_enum c = ch->value.e;
    // RGB -> R / G / B: CHANNEL_EXTRACT_U8_U24_POS0..2
    if (in->format == VX_DF_IMAGE_RGB) {
        if (c != VX_CHANNEL_R && c != VX_CHANNEL_G && c != VX_CHANNEL_B && c != VX_CHANNEL_0 && c != VX_CHANNEL_1 && c != VX_CHANNEL_2)
            return fail(n, VX_ERROR_INVALID_VALUE, "RGB channels are R, G, B");
        return set_output(n, 2, in->width, in->height, VX_DF_IMAGE_U8);
    }
    if (in->format != VX_DF_IMAGE_IYUV && in->format != VX_DF_IMAGE_NV12 && in->format != VX_DF_IMAGE_NV21 && in->format != VX_DF_IMAGE_YUV4)
        return fail(n, VX_ERROR_NOT_SUPPORTED, "channel extraction is provided for RGB and YUV images");
    if (c == VX_CHANNEL_Y || c == VX_CHANNEL_0) return set_output(n, 2, in->width, in->height, VX_DF_IMAGE_U8);
    if ((c == VX_CHANNEL_U || c == VX_CHANNEL_V) && in->format != VX_DF_IMAGE_RGB) {
        const uint32_t d = in->format == VX_DF_IMAGE_YUV4 ? 1 : 2;
        return set_output(n, 2, in->width / d, in->height / d, VX_DF_IMAGE_U8);
    }
    return fail(n, VX_ERROR_NOT_SUPPORTED, "channel extraction: Y, U and V of YUV images");
}

vx_status x_extract(_vx_node *n, hipStream_t st)
{
    _vx_image *in = img(n, 0), *o = img(n, 2);
    const vx_enum c = scl(n, 1)->value.e;
    if (in->format == VX_DF_IMAGE_RGB) {
        const uint32_t pos = (c == VX_CHANNEL_R || c == VX_CHANNEL_0) ? 0u : (c == VX_CHANNEL_G || c == VX_CHANNEL_1) ? 1u : 2u;
        return rc(n, mvx_ChannelExtract_U8_U24(st, o->width, o->height, P(o), S(o), P(in), S(in), pos));
    }
    // NV12 / NV21 chroma: CHANNEL_EXTRACT_U8_U16_POS0 / _POS1 of the interleaved plane (ago_drama_divide.cpp:342-414)
    if ((in->format == VX_DF_IMAGE_NV12 || in->format == VX_DF_IMAGE_NV21) && c != VX_CHANNEL_Y && c != VX_CHANNEL_0) {
        const uint32_t pos = (c == VX_CHANNEL_U) == (in->format == VX_DF_IMAGE_NV12) ? 0u : 1u;
        return rc(n, mvx_ChannelExtract_U8_U16(st, o->width, o->height, P(o), S(o), P(in, 1), S(in, 1), pos));
    }
    const int plane = c == VX_CHANNEL_U ? 1 : c == VX_CHANNEL_V ? 2 : 0;
    return rc(n, mvx_ChannelCopy_U8_U8(st, o->width, o->height, P(o), S(o), P(in, plane), S(in, plane)));
}

vx_status v_amd_copy(_vx_node *n)
{
    vx_status s = need_input(n, 1, VX_DF_IMAGE_U8);
    if (!s) s = set_output(n, 0, img(n, 1)->width, img(n, 1)->height, VX_DF_IMAGE_U8);
    return s;
}

vx_status x_amd_copy(_vx_node *n, hipStream_t st)
{
    _vx_image *o = img(n, 0), *in = img(n, 1);
    return rc(n, mvx_ChannelCopy_U8_U8(st, o->width, o->height, P(o), S(o), P(in), S(in)));
}

// ---------------------------------------------------------------------- ScaleImage (A10) ---------
// scale_image (ago_drama_divide.cpp:563-616): same size -> CHANNEL_COPY; otherwise by interpolation
// and the node's border mode: NEAREST_NEIGHBOR -> SCALE_IMAGE_U8_U8_NEAREST (any border); BILINEAR ->
// _BILINEAR (UNDEFINED), _BILINEAR_REPLICATE, _BILINEAR_CONSTANT (with the border value); AREA ->
// _AREA (any border).  The reference's optimizer then makes a same-size area scale a copy and one
// that does not shrink both dimensions nearest (ago_drama_remove.cpp:1270-1311); of the shrinking
// ratios only an exact 2x has a defined result there
// the sizes each ScaleImage kernel provides: kind 0 nearest, 1 bilinear, 2 bilinear replicate,
// 3 bilinear constant, 4 area
static vx_status scale_supported(_vx_node *n, const _vx_image *in, const _vx_image *o, int kind)
{
    if (in->width == o->width && in->height == o->height) return VX_SUCCESS;   // a plane copy
    if (kind == 4 && in->width > o->width && in->height > o->height) {
        const float xs = (float)((double)in->width / (double)o->width), ys = (float)((double)in->height / (double)o->height);
        if (!(xs == 2.0f && ys == 2.0f && o->width >= 8))
            return fail(n, VX_ERROR_NOT_SUPPORTED, "a shrinking AREA scale is provided for an exact 2x (output width >= 8) "
                                                   "only: the reference's generic area path has no defined result");
        return VX_SUCCESS;
    }
    if ((kind == 0 || kind == 4) && (in->width > 65535 || in->height > 65535))
        return fail(n, VX_ERROR_NOT_SUPPORTED, "nearest-neighbour sources above 65535 pixels (16-bit maps)");
    if ((kind == 2 || kind == 3) && in->width < 2)
        return fail(n, VX_ERROR_NOT_SUPPORTED, "bilinear border forms need a source at least 2 pixels wide");
    return VX_SUCCESS;
}

vx_status v_scale(_vx_node *n)
{
    vx_status s = need_input(n, 0, VX_DF_IMAGE_U8);
    if (s) return s;
    _vx_image *in = img(n, 0), *o = img(n, 1);
    _vx_scalar *type = scl(n, 2);
    if (!o || !type) return fail(n, VX_ERROR_INVALID_PARAMETERS, "arguments");
    const vx_enum t = type->value.e;
    if (t != VX_INTERPOLATION_NEAREST_NEIGHBOR && t != VX_INTERPOLATION_BILINEAR && t != VX_INTERPOLATION_AREA)
        return fail(n, VX_ERROR_INVALID_VALUE, "interpolation type");
    if (n->border.mode != VX_BORDER_UNDEFINED && n->border.mode != VX_BORDER_REPLICATE && n->border.mode != VX_BORDER_CONSTANT)
        return fail(n, VX_ERROR_NOT_SUPPORTED, "border mode");
    if (o->width == 0 || o->height == 0) return fail(n, VX_ERROR_INVALID_DIMENSION, "output size must be set");
    const int kind = t == VX_INTERPOLATION_NEAREST_NEIGHBOR ? 0 : t == VX_INTERPOLATION_AREA ? 4
                   : n->border.mode == VX_BORDER_REPLICATE ? 2 : n->border.mode == VX_BORDER_CONSTANT ? 3 : 1;
    s = scale_supported(n, in, o, kind);
    if (s) return s;
    return set_output(n, 1, o->width, o->height, VX_DF_IMAGE_U8);
}

vx_status x_scale(_vx_node *n, hipStream_t st)
{
    _vx_image *in = img(n, 0), *o = img(n, 1);
    if (in->width == o->width && in->height == o->height)
        return rc(n, mvx_ChannelCopy_U8_U8(st, o->width, o->height, P(o), S(o), P(in), S(in)));
    const vx_enum t = scl(n, 2)->value.e;
    if (t == VX_INTERPOLATION_NEAREST_NEIGHBOR)
        return rc(n, mvx_ScaleImage_U8_U8_Nearest(st, o->width, o->height, P(o), S(o), in->width, in->height, P(in), S(in)));
    if (t == VX_INTERPOLATION_AREA)
        return rc(n, mvx_ScaleImage_U8_U8_Area(st, o->width, o->height, P(o), S(o), in->width, in->height, P(in), S(in)));
    if (n->border.mode == VX_BORDER_REPLICATE)
        return rc(n, mvx_ScaleImage_U8_U8_Bilinear_Replicate(st, o->width, o->height, P(o), S(o), in->width, in->height,
                                                             P(in), S(in)));
    if (n->border.mode == VX_BORDER_CONSTANT)
        return rc(n, mvx_ScaleImage_U8_U8_Bilinear_Constant(st, o->width, o->height, P(o), S(o), in->width, in->height,
                                                            P(in), S(in), n->border.constant_value.U8));
    return rc(n, mvx_ScaleImage_U8_U8_Bilinear(st, o->width, o->height, P(o), S(o), in->width, in->height, P(in), S(in)));
}

vx_status v_amd_scale(_vx_node *n)
{
    vx_status s = need_input(n, 1, VX_DF_IMAGE_U8);
    if (s) return s;
    _vx_image *o = img(n, 0);
    if (!o || o->width == 0 || o->height == 0) return fail(n, VX_ERROR_INVALID_DIMENSION, "output size must be set");
    return set_output(n, 0, o->width, o->height, VX_DF_IMAGE_U8);
}

// AMD forms (out, in[, border scalar]): one ScaleImage kernel each
template <int KIND> vx_status x_amd_scale(_vx_node *n, hipStream_t st)
{
    _vx_image *o = img(n, 0), *in = img(n, 1);
    const uint32_t dw = o->width, dh = o->height, sw = in->width, sh = in->height;
    // every AMD ScaleImage kernel runs as a plane copy at the same size (ago_drama_remove.cpp:1283-1288)
    if (dw == sw && dh == sh) return rc(n, mvx_ChannelCopy_U8_U8(st, dw, dh, P(o), S(o), P(in), S(in)));
    switch (KIND) {
    case 0: return rc(n, mvx_ScaleImage_U8_U8_Nearest(st, dw, dh, P(o), S(o), sw, sh, P(in), S(in)));
    case 1: return rc(n, mvx_ScaleImage_U8_U8_Bilinear(st, dw, dh, P(o), S(o), sw, sh, P(in), S(in)));
    case 2: return rc(n, mvx_ScaleImage_U8_U8_Bilinear_Replicate(st, dw, dh, P(o), S(o), sw, sh, P(in), S(in)));
    case 3: return rc(n, mvx_ScaleImage_U8_U8_Bilinear_Constant(st, dw, dh, P(o), S(o), sw, sh, P(in), S(in), scl(n, 2)->value.u8));
    default: return rc(n, mvx_ScaleImage_U8_U8_Area(st, dw, dh, P(o), S(o), sw, sh, P(in), S(in)));
    }
}

template <int KIND> vx_status v_amd_scale_k(_vx_node *n)
{
    vx_status s = v_amd_scale(n);
    if (!s && KIND == 3 && (!scl(n, 2) || scl(n, 2)->dataType != VX_TYPE_UINT8))
        s = fail(n, VX_ERROR_INVALID_FORMAT, "border scalar must be UINT8");
    if (!s) s = scale_supported(n, img(n, 1), img(n, 0), KIND);
    return s;
}

// ---------------------------------------------------------------------- GaussianPyramid (A11) ----
// level 0 CHANNEL_COPY, each further level SCALE_GAUSSIAN_HALF_U8_U8_5x5 (ago_drama_divide.cpp:900-922)
vx_status v_pyramid(_vx_node *n)
{
    vx_status s = need_input(n, 0, VX_DF_IMAGE_U8);
    if (s) return s;
    if (!is_ref(n->params[1], VX_TYPE_PYRAMID)) return fail(n, VX_ERROR_INVALID_PARAMETERS, "output is not a pyramid");
    _vx_pyramid *p = (_vx_pyramid *)n->params[1];
    if (p->scale != VX_SCALE_PYRAMID_HALF) return fail(n, VX_ERROR_NOT_SUPPORTED, "only HALF pyramids are provided");
    if (p->format == VX_DF_IMAGE_VIRT) p->format = VX_DF_IMAGE_U8;
    if (p->format != VX_DF_IMAGE_U8) return fail(n, VX_ERROR_INVALID_FORMAT, "pyramid format");
    _vx_image *in = img(n, 0);
    if (p->width == 0 || p->height == 0) {
        if (!p->isVirtual) return fail(n, VX_ERROR_INVALID_DIMENSION, "pyramid size");
        p->width = in->width;
        p->height = in->height;
        if (!p->build_levels()) return fail(n, VX_ERROR_INVALID_FORMAT, "pyramid levels");
    }
    if (p->width != in->width || p->height != in->height) return fail(n, VX_ERROR_INVALID_DIMENSION, "pyramid size");
    for (size_t l = 1; l < p->images.size(); l++)
        if (p->images[l - 1]->width < 5 || p->images[l - 1]->height < 5 || p->images[l]->width < 3 || p->images[l]->height < 3)
            return fail(n, VX_ERROR_INVALID_DIMENSION, "pyramid level below the 5x5 filter's minimum size");
    return border_undefined(n);
}

vx_status x_pyramid(_vx_node *n, hipStream_t st)
{
    _vx_image *in = img(n, 0);
    _vx_pyramid *p = (_vx_pyramid *)n->params[1];
    _vx_image *l0 = p->images[0];
    vx_status s = rc(n, mvx_ChannelCopy_U8_U8(st, l0->width, l0->height, P(l0), S(l0), P(in), S(in)));
    for (size_t l = 1; l < p->images.size() && !s; l++) {
        _vx_image *a = p->images[l], *b = p->images[l - 1];
        s = rc(n, mvx_ScaleGaussianHalf_U8_U8_5x5(st, a->width, a->height, P(a), S(a), b->width, b->height, P(b), S(b)));
    }
    return s;
}

// the sizes each half-scale kernel provides: 3 -> ScaleGaussianHalf 3x3 (an output at least 2 rows high:
// the reference's dstHeight - 2 wraps below that), 5 -> ScaleGaussianHalf 5x5 (source >= 5 and output
// >= 3 in each dimension, ago_kernel_api.cpp:15428-15430), 1 -> ScaleImage nearest
static vx_status half_supported(_vx_node *n, const _vx_image *in, int ksize)
{
    const uint32_t dw = (in->width + 1) / 2, dh = (in->height + 1) / 2;
    if (ksize == 3 && dh < 2) return fail(n, VX_ERROR_INVALID_DIMENSION, "ScaleGaussianHalf 3x3 needs an output 2 rows high");
    if (ksize == 5 && (in->width < 5 || in->height < 5 || dw < 3 || dh < 3))
        return fail(n, VX_ERROR_INVALID_DIMENSION, "ScaleGaussianHalf 5x5 needs a source of at least 5x5");
    if (ksize == 1 && (in->width > 65535 || in->height > 65535))
        return fail(n, VX_ERROR_NOT_SUPPORTED, "nearest-neighbour sources above 65535 pixels (16-bit maps)");
    return VX_SUCCESS;
}

static vx_status run_half(_vx_node *n, hipStream_t st, _vx_image *o, _vx_image *in, int ksize)
{
    if (ksize == 1)
        return rc(n, mvx_ScaleImage_U8_U8_Nearest(st, o->width, o->height, P(o), S(o), in->width, in->height, P(in), S(in)));
    if (ksize == 3)
        return rc(n, mvx_ScaleGaussianHalf_U8_U8_3x3(st, o->width, o->height, P(o), S(o), in->width, in->height, P(in), S(in)));
    return rc(n, mvx_ScaleGaussianHalf_U8_U8_5x5(st, o->width, o->height, P(o), S(o), in->width, in->height, P(in), S(in)));
}

// AMD forms (out, in): KSIZE 3 or 5
template <int KSIZE> vx_status v_amd_half(_vx_node *n)
{
    vx_status s = need_input(n, 1, VX_DF_IMAGE_U8);
    if (s) return s;
    _vx_image *in = img(n, 1);
    s = half_supported(n, in, KSIZE);
    if (s) return s;
    return set_output(n, 0, (in->width + 1) / 2, (in->height + 1) / 2, VX_DF_IMAGE_U8);
}

template <int KSIZE> vx_status x_amd_half(_vx_node *n, hipStream_t st) { return run_half(n, st, img(n, 0), img(n, 1), KSIZE); }

// halfscale_gaussian (ago_drama_divide.cpp:1471-1488): kernel size 1 -> SCALE_IMAGE_U8_U8_NEAREST,
// 3 -> SCALE_GAUSSIAN_HALF_U8_U8_3x3, 5 -> SCALE_GAUSSIAN_HALF_U8_U8_5x5; output (w+1)/2 x (h+1)/2
vx_status v_halfscale(_vx_node *n)
{
    vx_status s = need_input(n, 0, VX_DF_IMAGE_U8);
    if (s) return s;
    _vx_scalar *k = scl(n, 2);
    if (!k || k->dataType != VX_TYPE_INT32) return fail(n, VX_ERROR_INVALID_PARAMETERS, "kernel_size is not an INT32 scalar");
    const int ks = k->value.i32;
    if (ks != 1 && ks != 3 && ks != 5) return fail(n, VX_ERROR_INVALID_VALUE, "kernel_size must be 1, 3 or 5");
    _vx_image *in = img(n, 0);
    s = half_supported(n, in, ks);
    if (s) return s;
    return set_output(n, 1, (in->width + 1) / 2, (in->height + 1) / 2, VX_DF_IMAGE_U8);
}

vx_status x_halfscale(_vx_node *n, hipStream_t st) { return run_half(n, st, img(n, 1), img(n, 0), scl(n, 2)->value.i32); }

// ---------------------------------------------------------------------- Canny (A12) --------------
// CANNY_SOBEL_U16_U8_3x3_L1NORM -> CANNY_SUPP_THRESHOLD_U8XY_U16_3x3 -> CANNY_EDGE_TRACE_U8_U8XY with a
// virtual stack of w*h items (ago_drama_divide.cpp:1618-1701).  Workspace: the U16 gradient plane and
// the stack (the trace's union-find labels).
vx_status v_canny(_vx_node *n)
{
    vx_status s = need_input(n, 0, VX_DF_IMAGE_U8);
    if (s) return s;
    if (!is_ref(n->params[1], VX_TYPE_THRESHOLD)) return fail(n, VX_ERROR_INVALID_PARAMETERS, "hysteresis is not a threshold");
    _vx_threshold *t = (_vx_threshold *)n->params[1];
    if (t->threshType != VX_THRESHOLD_TYPE_RANGE) return fail(n, VX_ERROR_INVALID_TYPE, "hysteresis threshold must be RANGE");
    _vx_scalar *g = scl(n, 2), *norm = scl(n, 3);
    if (!g || !norm) return fail(n, VX_ERROR_INVALID_PARAMETERS, "gradient size / norm");
    if (g->value.i32 != 3 && g->value.i32 != 5 && g->value.i32 != 7) return fail(n, VX_ERROR_INVALID_VALUE, "gradient size");
    if (norm->value.e != VX_NORM_L1 && norm->value.e != VX_NORM_L2) return fail(n, VX_ERROR_INVALID_VALUE, "norm type");
    _vx_image *in = img(n, 0);
    if (in->width < 3 || in->height < 3) return fail(n, VX_ERROR_INVALID_DIMENSION, "image below 3x3");
    if ((s = border_undefined(n))) return s;
    return set_output(n, 4, in->width, in->height, VX_DF_IMAGE_U8);
}

size_t canny_mag_bytes(uint32_t w, uint32_t h) { return align256((size_t)align16(2 * w) * h); }

vx_status i_canny(_vx_node *n)
{
    _vx_image *in = img(n, 0);
    n->localSize = canny_mag_bytes(in->width, in->height) + (size_t)4 * in->width * in->height + 256;
    return VX_SUCCESS;
}

vx_status x_canny(_vx_node *n, hipStream_t st)
{
    _vx_image *in = img(n, 0), *o = img(n, 4);
    _vx_threshold *t = (_vx_threshold *)n->params[1];
    const uint32_t w = in->width, h = in->height, magStride = align16(2 * w);
    uint16_t *mag = (uint16_t *)n->local.ptr();
    // the trace's stack at a 256-byte boundary (its tables then start on whole cache lines)
    void *stack = (void *)(((uintptr_t)(n->local.ptr() + canny_mag_bytes(w, h)) + 255u) & ~(uintptr_t)255u);
    const int g = scl(n, 2)->value.i32;
    const bool l2 = scl(n, 3)->value.e == VX_NORM_L2;
    // gradient 7 runs CANNY_SUPP_THRESHOLD_U8XY_U16_7x7, which divides the threshold object's values by
    // 4 in place on every run before using them (ago_kernel_api.cpp:18033-18035)
    if (g == 7) {
        t->lower.S32 /= 4;
        t->upper.S32 /= 4;
    }
    // thresholds as the reference reads them: the low 32 bits of the pixel values (ago_kernel_api.cpp:17969)
    const uint32_t lower = t->lower.U32, upper = t->upper.U32;
    const uint8_t *ip = P(in);
    const uint32_t is = S(in);
    vx_status s = rc(n, g == 3 ? (l2 ? mvx_CannySobel_U16_U8_3x3_L2NORM(st, w, h, mag, magStride, ip, is)
                                     : mvx_CannySobel_U16_U8_3x3_L1NORM(st, w, h, mag, magStride, ip, is))
                     : g == 5 ? (l2 ? mvx_CannySobel_U16_U8_5x5_L2NORM(st, w, h, mag, magStride, ip, is)
                                    : mvx_CannySobel_U16_U8_5x5_L1NORM(st, w, h, mag, magStride, ip, is))
                              : (l2 ? mvx_CannySobel_U16_U8_7x7_L2NORM(st, w, h, mag, magStride, ip, is)
                                    : mvx_CannySobel_U16_U8_7x7_L1NORM(st, w, h, mag, magStride, ip, is)));
    if (!s) s = rc(n, mvx_CannySuppThreshold_U8XY_U16_3x3(st, w, h, P(o), S(o), mag, magStride, stack, w * h, lower, upper));
    if (!s) s = rc(n, mvx_CannyEdgeTrace_U8_U8XY(st, w, h, P(o), S(o), stack, w * h));
    return s;
}

vx_status v_amd_canny_sobel(_vx_node *n)
{
    vx_status s = need_input(n, 1, VX_DF_IMAGE_U8);
    if (!s) s = set_output(n, 0, img(n, 1)->width, img(n, 1)->height, VX_DF_IMAGE_U16);
    return s;
}

// AMD forms CannySobel_U16_U8_<g>x<g>_<norm> (out U16, in U8)
template <int G, bool L2> vx_status x_amd_canny_sobel(_vx_node *n, hipStream_t st)
{
    _vx_image *o = img(n, 0), *in = img(n, 1);
    uint16_t *d = (uint16_t *)P(o);
    const uint32_t w = o->width, h = o->height, ds = S(o), ss = S(in);
    const uint8_t *sp = P(in);
    if (G == 3) return rc(n, L2 ? mvx_CannySobel_U16_U8_3x3_L2NORM(st, w, h, d, ds, sp, ss) : mvx_CannySobel_U16_U8_3x3_L1NORM(st, w, h, d, ds, sp, ss));
    if (G == 5) return rc(n, L2 ? mvx_CannySobel_U16_U8_5x5_L2NORM(st, w, h, d, ds, sp, ss) : mvx_CannySobel_U16_U8_5x5_L1NORM(st, w, h, d, ds, sp, ss));
    return rc(n, L2 ? mvx_CannySobel_U16_U8_7x7_L2NORM(st, w, h, d, ds, sp, ss) : mvx_CannySobel_U16_U8_7x7_L1NORM(st, w, h, d, ds, sp, ss));
}

// ---------------------------------------------------------------------- Harris (A13) -------------
// HARRIS_SOBEL_HG3_U8_3x3 -> HARRIS_SCORE_HVC_HG3_3x3 -> NON_MAX_SUPP_XY_ANY_3x3 ->
// HARRIS_MERGE_SORT_AND_PICK_XY_XYS (ago_drama_divide.cpp:1703-1787)
struct HarrisWs {
    size_t hg3, hvc, xys, cnt, nms, pick, total;
    uint32_t maxList;
    size_t nmsSize, pickSize;
};

HarrisWs harris_ws(uint32_t w, uint32_t h)
{
    HarrisWs L;
    // non-maxima are never 8-adjacent: at most one per 2x2 block of the interior
    L.maxList = std::max(1u, ((w - 1) / 2) * ((h - 1) / 2));
    L.nmsSize = mvx_NonMaxSupp_XY_ANY_3x3_workspace_size(w, h);
    L.pickSize = mvx_HarrisMergeSortAndPick_XY_XYS_workspace_size(L.maxList);
    L.hg3 = 0;
    L.hvc = align256(L.hg3 + (size_t)align16(12 * w) * h);
    L.xys = align256(L.hvc + (size_t)align16(4 * w) * h);
    L.cnt = align256(L.xys + (size_t)8 * L.maxList);
    L.nms = align256(L.cnt + 4);
    L.pick = align256(L.nms + L.nmsSize);
    L.total = align256(L.pick + L.pickSize);
    return L;
}

// (255 * 2^(g-1) * b)^4 in float, left to right (ago_kernel_api.cpp:16941-16942)
float harris_norm(int gradient, int block)
{
    const float f = 255.0f * (float)(1 << (gradient - 1)) * (float)block;
    return ((f * f) * f) * f;
}

vx_status v_harris(_vx_node *n)
{
    vx_status s = need_input(n, 0, VX_DF_IMAGE_U8);
    if (s) return s;
    _vx_scalar *thr = scl(n, 1), *dist = scl(n, 2), *k = scl(n, 3), *g = scl(n, 4), *b = scl(n, 5);
    if (!thr || !dist || !k || !g || !b) return fail(n, VX_ERROR_INVALID_PARAMETERS, "scalars");
    if (thr->dataType != VX_TYPE_FLOAT32 || dist->dataType != VX_TYPE_FLOAT32 || k->dataType != VX_TYPE_FLOAT32)
        return fail(n, VX_ERROR_INVALID_TYPE, "strength / distance / sensitivity must be FLOAT32");
    for (int v : {g->value.i32, b->value.i32})
        if (v != 3 && v != 5 && v != 7) return fail(n, VX_ERROR_INVALID_VALUE, "gradient and block sizes must be 3, 5 or 7");
    if (!is_ref(n->params[6], VX_TYPE_ARRAY) || ((_vx_array *)n->params[6])->itemType != VX_TYPE_KEYPOINT)
        return fail(n, VX_ERROR_INVALID_PARAMETERS, "corners must be a KEYPOINT array");
    _vx_array *corners = (_vx_array *)n->params[6];
    if (!corners->capacity) return fail(n, VX_ERROR_INVALID_DIMENSION, "corner array capacity");
    if (n->params[7] && (!scl(n, 7) || scl(n, 7)->dataType != VX_TYPE_SIZE))
        return fail(n, VX_ERROR_INVALID_TYPE, "num_corners must be a SIZE scalar");
    _vx_image *in = img(n, 0);
    const uint32_t minSide = (uint32_t)std::max(3, std::max(g->value.i32, b->value.i32));
    if (in->width < minSide || in->height < minSide || in->width >= 32768 || in->height >= 32768)
        return fail(n, VX_ERROR_INVALID_DIMENSION, "image size");
    return VX_SUCCESS;
}

vx_status i_harris(_vx_node *n)
{
    _vx_image *in = img(n, 0);
    n->localSize = harris_ws(in->width, in->height).total;
    return VX_SUCCESS;
}

vx_status x_harris(_vx_node *n, hipStream_t st)
{
    _vx_image *in = img(n, 0);
    const uint32_t w = in->width, h = in->height;
    const HarrisWs L = harris_ws(w, h);
    uint8_t *ws = n->local.ptr();
    _vx_array *corners = (_vx_array *)n->params[6];
    _vx_scalar *num = scl(n, 7);
    const float thr = scl(n, 1)->value.f32, dist = scl(n, 2)->value.f32, k = scl(n, 3)->value.f32;
    float *hg3 = (float *)(ws + L.hg3), *hvc = (float *)(ws + L.hvc);
    uint32_t *cnt = (uint32_t *)(ws + L.cnt);
    // HARRIS_SOBEL_HG3_U8_<g>x<g> -> HARRIS_SCORE_HVC_HG3_<b>x<b> (ago_drama_divide.cpp:1703-1787)
    const int g = scl(n, 4)->value.i32, b = scl(n, 5)->value.i32;
    const uint32_t hs = align16(12 * w), vs = align16(4 * w);
    vx_status s = rc(n, g == 3 ? mvx_HarrisSobel_HG3_U8_3x3(st, w, h, hg3, hs, P(in), S(in))
                     : g == 5 ? mvx_HarrisSobel_HG3_U8_5x5(st, w, h, hg3, hs, P(in), S(in))
                              : mvx_HarrisSobel_HG3_U8_7x7(st, w, h, hg3, hs, P(in), S(in)));
    const float norm = harris_norm(g, b);
    if (!s) s = rc(n, b == 3 ? mvx_HarrisScore_HVC_HG3_3x3(st, w, h, hvc, vs, hg3, hs, k, thr, norm)
                      : b == 5 ? mvx_HarrisScore_HVC_HG3_5x5(st, w, h, hvc, vs, hg3, hs, k, thr, norm)
                               : mvx_HarrisScore_HVC_HG3_7x7(st, w, h, hvc, vs, hg3, hs, k, thr, norm));
    if (!s) s = rc(n, mvx_NonMaxSupp_XY_ANY_3x3(st, L.maxList, ws + L.xys, cnt, w, h, hvc, align16(4 * w), ws + L.nms, L.nmsSize));
    // the count scalar lives in the node workspace's tail when the caller gave none
    uint64_t *count = num ? (uint64_t *)num->dev.ptr() : (uint64_t *)(ws + L.cnt + 8);
    if (!s) s = rc(n, mvx_HarrisMergeSortAndPick_XY_XYS(st, (uint32_t)corners->capacity, corners->items_ptr(), corners->count_ptr(),
                                                        count, ws + L.xys, cnt, L.maxList, dist, w, h, ws + L.pick, L.pickSize));
    return s;
}

// the pick workspace must be prepared once (hash slots empty)
vx_status i_harris_prepare(_vx_node *n, hipStream_t st)
{
    _vx_image *in = img(n, 0);
    const HarrisWs L = harris_ws(in->width, in->height);
    return rc(n, mvx_HarrisMergeSortAndPick_workspace_init(st, L.maxList, n->local.ptr() + L.pick, L.pickSize));
}

// ---------------------------------------------------------------------- FAST corners -------------
// fast_corners -> FAST_CORNERS_XY_U8_SUPRESSION, or _NOSUPRESSION when nonmax_suppression is false
// (ago_drama_divide.cpp:1789-1808); array count = min(corners, capacity), num_corners = all corners
// (ago_kernel_api.cpp:16597-16604)
static vx_status fast_check(_vx_node *n, int img_i, int thr_i, int arr_i, int num_i)
{
    vx_status s = need_input(n, img_i, VX_DF_IMAGE_U8);
    if (s) return s;
    _vx_scalar *thr = scl(n, thr_i);
    if (!thr || thr->dataType != VX_TYPE_FLOAT32) return fail(n, VX_ERROR_INVALID_TYPE, "strength_thresh must be FLOAT32");
    if (!is_ref(n->params[arr_i], VX_TYPE_ARRAY) || ((_vx_array *)n->params[arr_i])->itemType != VX_TYPE_KEYPOINT)
        return fail(n, VX_ERROR_INVALID_PARAMETERS, "corners must be a KEYPOINT array");
    if (!((_vx_array *)n->params[arr_i])->capacity) return fail(n, VX_ERROR_INVALID_DIMENSION, "corner array capacity");
    if (n->params[num_i] && (!scl(n, num_i) || scl(n, num_i)->dataType != VX_TYPE_SIZE))
        return fail(n, VX_ERROR_INVALID_TYPE, "num_corners must be a SIZE scalar");
    _vx_image *in = img(n, img_i);
    if (in->width < 7 || in->height < 7 || in->width > 32767 || in->height > 32767)
        return fail(n, VX_ERROR_INVALID_DIMENSION, "FAST needs an image of at least 7x7");
    return VX_SUCCESS;
}

vx_status v_fast(_vx_node *n)
{
    vx_status s = fast_check(n, 0, 1, 3, 4);
    if (!s && (!scl(n, 2) || scl(n, 2)->dataType != VX_TYPE_BOOL)) s = fail(n, VX_ERROR_INVALID_TYPE, "nonmax_suppression must be BOOL");
    return s;
}

vx_status i_fast(_vx_node *n)
{
    _vx_image *in = img(n, 0);
    n->localSize = mvx_FastCorners_workspace_size(in->width, in->height) + 256;
    return VX_SUCCESS;
}

static vx_status run_fast(_vx_node *n, hipStream_t st, _vx_image *in, float thr, bool nonmax, _vx_array *corners, _vx_scalar *num)
{
    const size_t wsz = mvx_FastCorners_workspace_size(in->width, in->height);
    uint8_t *ws = n->local.ptr();
    // the count scalar lives in the node workspace's tail when the caller gave none
    uint64_t *count = num ? (uint64_t *)num->dev.ptr() : (uint64_t *)(ws + ((wsz + 7) & ~(size_t)7));
    return rc(n, (nonmax ? mvx_FastCorners_XY_U8_Supression : mvx_FastCorners_XY_U8_NoSupression)(
                     st, (uint32_t)corners->capacity, corners->items_ptr(), corners->count_ptr(), count, in->width, in->height, P(in), S(in),
                     thr, ws, wsz));
}

vx_status x_fast(_vx_node *n, hipStream_t st)
{
    return run_fast(n, st, img(n, 0), scl(n, 1)->value.f32, scl(n, 2)->value.b != vx_false_e, (_vx_array *)n->params[3], scl(n, 4));
}

// AMD forms: (corners out, num_corners out, image, strength_thresh)
vx_status v_amd_fast(_vx_node *n) { return fast_check(n, 2, 3, 0, 1); }

vx_status i_amd_fast(_vx_node *n)
{
    _vx_image *in = img(n, 2);
    n->localSize = mvx_FastCorners_workspace_size(in->width, in->height) + 256;
    return VX_SUCCESS;
}

template <bool NONMAX> vx_status x_amd_fast(_vx_node *n, hipStream_t st)
{
    return run_fast(n, st, img(n, 2), scl(n, 3)->value.f32, NONMAX, (_vx_array *)n->params[0], scl(n, 1));
}

// ---------------------------------------------------------------------- warps (A14, A15) ---------
// warp_affine / warp_perspective bilinear -> WARP_*_U8_U8_BILINEAR (border UNDEFINED) or
// WARP_*_U8_U8_BILINEAR_CONSTANT with the border value (ago_drama_divide.cpp:1542-1616); the matrix
// goes to the kernel as stored (2x3 or 3x3 FLOAT32, ago_kernel_api.cpp:19143-19275).
vx_status check_matrix(_vx_node *n, int i, vx_size cols)
{
    if (!is_ref(n->params[i], VX_TYPE_MATRIX)) return fail(n, VX_ERROR_INVALID_PARAMETERS, "matrix");
    _vx_matrix *m = (_vx_matrix *)n->params[i];
    if (m->dataType != VX_TYPE_FLOAT32) return fail(n, VX_ERROR_INVALID_TYPE, "matrix must be FLOAT32");
    if (m->columns != cols || m->rows != 3) return fail(n, VX_ERROR_INVALID_DIMENSION, "matrix dimensions");
    return VX_SUCCESS;
}

vx_status fixed_size_output(_vx_node *n, int i)
{
    _vx_image *o = img(n, i);
    if (!o || o->width == 0 || o->height == 0) return fail(n, VX_ERROR_INVALID_DIMENSION, "output size must be set");
    return set_output(n, i, o->width, o->height, VX_DF_IMAGE_U8);
}

vx_status interp_and_border(_vx_node *n, int interp)
{
    _vx_scalar *t = scl(n, interp);
    if (!t) return fail(n, VX_ERROR_INVALID_PARAMETERS, "interpolation");
    if (t->value.e != VX_INTERPOLATION_BILINEAR && t->value.e != VX_INTERPOLATION_NEAREST_NEIGHBOR)
        return fail(n, VX_ERROR_INVALID_VALUE, "interpolation must be NEAREST_NEIGHBOR or BILINEAR");
    if (n->border.mode != VX_BORDER_UNDEFINED && n->border.mode != VX_BORDER_CONSTANT)
        return fail(n, VX_ERROR_NOT_SUPPORTED, "border mode must be UNDEFINED or CONSTANT");
    return VX_SUCCESS;
}

template <bool PERSP> vx_status v_warp(_vx_node *n)
{
    vx_status s = need_input(n, 0, VX_DF_IMAGE_U8);
    if (!s) s = check_matrix(n, 1, PERSP ? 3 : 2);
    if (!s) s = interp_and_border(n, 2);
    if (!s) s = fixed_size_output(n, 3);
    return s;
}

// warp_affine / warp_perspective -> WARP_*_U8_U8_{NEAREST,BILINEAR}[_CONSTANT] by interpolation and node
// border (ago_drama_divide.cpp:1542-1616).  This engine keeps no valid-region rectangles: the nearest
// constant affine form is given the whole source as its valid region.
template <bool PERSP>
vx_status run_warp(_vx_node *n, hipStream_t st, _vx_image *o, _vx_image *in, _vx_matrix *m, bool constant, uint8_t border,
                   bool nearest)
{
    const float *mat = (const float *)m->data.data();
    if (nearest) {
        if (PERSP)
            return rc(n, constant ? mvx_WarpPerspective_U8_U8_Nearest_Constant(st, o->width, o->height, P(o), S(o), in->width,
                                                                               in->height, P(in), S(in), mat, border)
                                  : mvx_WarpPerspective_U8_U8_Nearest(st, o->width, o->height, P(o), S(o), in->width, in->height,
                                                                      P(in), S(in), mat));
        return rc(n, constant ? mvx_WarpAffine_U8_U8_Nearest_Constant(st, o->width, o->height, P(o), S(o), in->width, in->height,
                                                                      P(in), S(in), mat, border, nullptr)
                              : mvx_WarpAffine_U8_U8_Nearest(st, o->width, o->height, P(o), S(o), in->width, in->height, P(in),
                                                             S(in), mat));
    }
    if (PERSP)
        return rc(n, constant ? mvx_WarpPerspective_U8_U8_Bilinear_Constant(st, o->width, o->height, P(o), S(o), in->width, in->height,
                                                                            P(in), S(in), mat, border)
                              : mvx_WarpPerspective_U8_U8_Bilinear(st, o->width, o->height, P(o), S(o), in->width, in->height, P(in),
                                                                   S(in), mat));
    return rc(n, constant ? mvx_WarpAffine_U8_U8_Bilinear_Constant(st, o->width, o->height, P(o), S(o), in->width, in->height, P(in),
                                                                   S(in), mat, border)
                          : mvx_WarpAffine_U8_U8_Bilinear(st, o->width, o->height, P(o), S(o), in->width, in->height, P(in), S(in), mat));
}

template <bool PERSP> vx_status x_warp(_vx_node *n, hipStream_t st)
{
    const bool constant = n->border.mode == VX_BORDER_CONSTANT;
    return run_warp<PERSP>(n, st, img(n, 3), img(n, 0), (_vx_matrix *)n->params[1], constant, n->border.constant_value.U8,
                           scl(n, 2)->value.e == VX_INTERPOLATION_NEAREST_NEIGHBOR);
}

// AMD forms: (out, in, matrix[, border scalar])
template <bool PERSP, bool CONST> vx_status v_amd_warp(_vx_node *n)
{
    vx_status s = need_input(n, 1, VX_DF_IMAGE_U8);
    if (!s) s = check_matrix(n, 2, PERSP ? 3 : 2);
    if (!s && CONST && (!scl(n, 3) || scl(n, 3)->dataType != VX_TYPE_UINT8)) s = fail(n, VX_ERROR_INVALID_FORMAT, "border scalar must be UINT8");
    if (!s) s = fixed_size_output(n, 0);
    return s;
}

template <bool PERSP, bool CONST, bool NN> vx_status x_amd_warp(_vx_node *n, hipStream_t st)
{
    return run_warp<PERSP>(n, st, img(n, 0), img(n, 1), (_vx_matrix *)n->params[2], CONST, CONST ? scl(n, 3)->value.u8 : 0, NN);
}

// ---------------------------------------------------------------------- Remap (A16) --------------
// remap -> REMAP_U8_U8_{NEAREST,BILINEAR}[_CONSTANT] (ago_drama_divide.cpp:1490-1540) over
// the fixed-point table vxSetRemapPoint builds (here: built on the GPU from the float table)
vx_status check_remap(_vx_node *n, int i, _vx_image *in)
{
    if (!is_ref(n->params[i], VX_TYPE_REMAP)) return fail(n, VX_ERROR_INVALID_PARAMETERS, "table is not a remap");
    _vx_remap *r = (_vx_remap *)n->params[i];
    if (r->srcWidth != in->width || r->srcHeight != in->height) return fail(n, VX_ERROR_INVALID_DIMENSION, "remap source size");
    if (r->srcWidth >= 8192 || r->srcHeight >= 8192)
        return fail(n, VX_ERROR_NOT_SUPPORTED, "remap sources of 8192 pixels or more (fewer fractional bits) are not provided");
    return VX_SUCCESS;
}

vx_status v_remap(_vx_node *n)
{
    vx_status s = need_input(n, 0, VX_DF_IMAGE_U8);
    if (!s) s = check_remap(n, 1, img(n, 0));
    if (!s) s = interp_and_border(n, 2);
    if (s) return s;
    _vx_remap *r = (_vx_remap *)n->params[1];
    return set_output(n, 3, r->dstWidth, r->dstHeight, VX_DF_IMAGE_U8);
}

vx_status run_remap(_vx_node *n, hipStream_t st, _vx_image *o, _vx_image *in, _vx_remap *r, bool constant, uint8_t border,
                    bool nearest)
{
    vx_status s = r->sync(st);
    if (s) return fail(n, s, "remap table upload");
    const uint32_t srcBytes = S(in) * in->height;
    if (nearest)
        return rc(n, constant ? mvx_Remap_U8_U8_Nearest_Constant(st, o->width, o->height, P(o), S(o), in->width, in->height, P(in),
                                                                 S(in), srcBytes, r->devFixed.ptr(), r->dstWidth * 4, border)
                              : mvx_Remap_U8_U8_Nearest(st, o->width, o->height, P(o), S(o), in->width, in->height, P(in), S(in),
                                                        srcBytes, r->devFixed.ptr(), r->dstWidth * 4));
    return rc(n, constant ? mvx_Remap_U8_U8_Bilinear_Constant(st, o->width, o->height, P(o), S(o), in->width, in->height, P(in), S(in),
                                                              r->devFixed.ptr(), r->dstWidth * 4, border)
                          : mvx_Remap_U8_U8_Bilinear(st, o->width, o->height, P(o), S(o), in->width, in->height, P(in), S(in),
                                                     srcBytes, r->devFixed.ptr(), r->dstWidth * 4));
}

vx_status x_remap(_vx_node *n, hipStream_t st)
{
    return run_remap(n, st, img(n, 3), img(n, 0), (_vx_remap *)n->params[1], n->border.mode == VX_BORDER_CONSTANT,
                     n->border.constant_value.U8, scl(n, 2)->value.e == VX_INTERPOLATION_NEAREST_NEIGHBOR);
}

template <bool CONST> vx_status v_amd_remap(_vx_node *n)
{
    vx_status s = need_input(n, 1, VX_DF_IMAGE_U8);
    if (!s) s = check_remap(n, 2, img(n, 1));
    if (!s && CONST && (!scl(n, 3) || scl(n, 3)->dataType != VX_TYPE_UINT8)) s = fail(n, VX_ERROR_INVALID_FORMAT, "border scalar must be UINT8");
    if (s) return s;
    _vx_remap *r = (_vx_remap *)n->params[2];
    return set_output(n, 0, r->dstWidth, r->dstHeight, VX_DF_IMAGE_U8);
}

template <bool CONST, bool NN> vx_status x_amd_remap(_vx_node *n, hipStream_t st)
{
    return run_remap(n, st, img(n, 0), img(n, 1), (_vx_remap *)n->params[2], CONST, CONST ? scl(n, 3)->value.u8 : 0, NN);
}

// ---------------------------------------------------------------------- OpticalFlowPyrLK (A17) --
// optical_flow_pyr_lk -> OPTICAL_FLOW_PYR_LK_XY_XY (ago_drama_divide.cpp:1810-1891), validation of
// ago_kernel_api.cpp:287-317 and :20617-20630; the node workspace keeps the tracker's scratch state
// from run to run as the reference node's local data does (ago_kernel_api.cpp:20633-20640).
vx_status v_lk(_vx_node *n)
{
    for (int i : {0, 1})
        if (!is_ref(n->params[i], VX_TYPE_PYRAMID)) return fail(n, VX_ERROR_INVALID_PARAMETERS, "pyramids");
    for (int i : {2, 3, 4})
        if (!is_ref(n->params[i], VX_TYPE_ARRAY)) return fail(n, VX_ERROR_INVALID_PARAMETERS, "keypoint arrays");
    _vx_pyramid *po = (_vx_pyramid *)n->params[0], *pn = (_vx_pyramid *)n->params[1];
    _vx_array *ao = (_vx_array *)n->params[2], *ae = (_vx_array *)n->params[3], *an = (_vx_array *)n->params[4];
    _vx_scalar *term = scl(n, 5), *eps = scl(n, 6), *iters = scl(n, 7), *use = scl(n, 8), *win = scl(n, 9);
    if (!term || !eps || !iters || !use || !win) return fail(n, VX_ERROR_INVALID_PARAMETERS, "scalars");
    if (ao->itemType != VX_TYPE_KEYPOINT || ae->itemType != VX_TYPE_KEYPOINT || term->dataType != VX_TYPE_ENUM ||
        eps->dataType != VX_TYPE_FLOAT32 || iters->dataType != VX_TYPE_UINT32 || use->dataType != VX_TYPE_BOOL ||
        win->dataType != VX_TYPE_SIZE)
        return fail(n, VX_ERROR_INVALID_TYPE, "argument types");
    if (po->format != VX_DF_IMAGE_U8 || pn->format != VX_DF_IMAGE_U8) return fail(n, VX_ERROR_INVALID_FORMAT, "pyramid format");
    if (!po->width || !po->height || po->width != pn->width || po->height != pn->height || !ao->capacity ||
        ao->capacity != ae->capacity)
        return fail(n, VX_ERROR_INVALID_DIMENSION, "pyramid sizes / array capacities");
    if (term->value.e != VX_TERM_CRITERIA_ITERATIONS && term->value.e != VX_TERM_CRITERIA_EPSILON && term->value.e != VX_TERM_CRITERIA_BOTH)
        return fail(n, VX_ERROR_INVALID_VALUE, "termination");
    if (po->scale != pn->scale || po->levels != pn->levels) return fail(n, VX_ERROR_INVALID_VALUE, "pyramid scale / levels");
    if (win->value.sz < 3 || win->value.sz > 15) return fail(n, VX_ERROR_INVALID_VALUE, "window dimension");
    for (_vx_image *l : po->images)
        if (l->width < 8 || l->height < 3) return fail(n, VX_ERROR_NOT_SUPPORTED, "pyramid levels below 8x3 pixels");
    if (an->itemType != VX_TYPE_KEYPOINT) return fail(n, VX_ERROR_INVALID_TYPE, "new points must be KEYPOINT");
    if (an->isVirtual && !an->capacity) an->capacity = ao->capacity;
    if (an->capacity < ao->capacity) return fail(n, VX_ERROR_INVALID_DIMENSION, "new points capacity");
    return VX_SUCCESS;
}

vx_status i_lk(_vx_node *n)
{
    _vx_pyramid *po = (_vx_pyramid *)n->params[0];
    n->localSize = mvx_OpticalFlowPyrLK_workspace_size(po->width, po->height, (uint32_t)((_vx_array *)n->params[2])->capacity);
    return VX_SUCCESS;
}

vx_status x_lk(_vx_node *n, hipStream_t st)
{
    _vx_pyramid *po = (_vx_pyramid *)n->params[0], *pn = (_vx_pyramid *)n->params[1];
    _vx_array *ao = (_vx_array *)n->params[2], *ae = (_vx_array *)n->params[3], *an = (_vx_array *)n->params[4];
    // the reference refuses differing point counts (ago_kernel_api.cpp:20602); host counts are
    // exact for arrays no node of this graph writes
    auto hostKnown = [&](_vx_array *a) {
        for (_vx_array *w : n->graph->writtenArrays)
            if (w == a) return false;
        return true;
    };
    if (hostKnown(ao) && hostKnown(ae) && ao->numItems != ae->numItems) return fail(n, VX_ERROR_INVALID_DIMENSION, "point counts differ");
    std::vector<mvx_image_level_t> lo(po->levels), ln(pn->levels);
    for (size_t l = 0; l < po->levels; l++) {
        lo[l] = mvx_image_level_t{P(po->images[l]), po->images[l]->width, po->images[l]->height, S(po->images[l])};
        ln[l] = mvx_image_level_t{P(pn->images[l]), pn->images[l]->width, pn->images[l]->height, S(pn->images[l])};
    }
    return rc(n, mvx_OpticalFlowPyrLK_XY_XY(st, an->items_ptr(), an->count_ptr(), lo.data(), ln.data(), (uint32_t)po->levels, po->scale,
                                            ao->items_ptr(), ao->count_ptr(), ae->items_ptr(), (uint32_t)ao->capacity,
                                            scl(n, 5)->value.e, scl(n, 6)->value.f32, scl(n, 7)->value.u32, scl(n, 8)->value.b ? 1 : 0,
                                            (uint32_t)scl(n, 9)->value.sz, n->local.ptr(), n->localSize));
}

// ---------------------------------------------------------------------- valid rectangles --------
// The reference's valid_rect_callback commands (ago_kernel_api.cpp, one per AMD kernel, e.g. Gaussian
// :15317-15326, Magnitude :10357-10364, ScaleImage :19897-19908, ScaleGaussianHalf 5x5 :15444-15453),
// composed per node the way the dividers split it.  Kernels without one (warps, remap, corner and
// tracking nodes, the Canny trace) leave their outputs' rectangles as they are.
vx_rectangle_t rect_of(_vx_image *im)
{
    vx_rectangle_t r;
    vxGetValidRegionImage(im, &r);
    return r;
}

void set_rect(_vx_image *im, vx_rectangle_t r)
{
    if (!im) return;
    im->validRect = r;
    im->validRectSet = true;
}

// start + k clamped to the output size, end - k clamped at 0 (the 3x3 / 5x5 / 7x7 stencil rules)
vx_rectangle_t shrink(vx_rectangle_t r, int k, _vx_image *out)
{
    return vx_rectangle_t{std::min(r.start_x + k, out->width), std::min(r.start_y + k, out->height),
                          (vx_uint32)std::max((int)r.end_x - k, 0), (vx_uint32)std::max((int)r.end_y - k, 0)};
}

// ScaleGaussianHalf 5x5 (:15444-15453); with `inner` the 3x3 form, one more pixel in (:15510-15518)
vx_rectangle_t half_rect(vx_rectangle_t r, _vx_image *out, bool inner)
{
    const int k = inner ? 1 : 0;
    return vx_rectangle_t{std::min(((r.start_x + 1) >> 1) + k, out->width), std::min(((r.start_y + 1) >> 1) + k, out->height),
                          (vx_uint32)std::max((int)((r.end_x + 1) >> 1) - k, 0), (vx_uint32)std::max((int)((r.end_y + 1) >> 1) - k, 0)};
}

// The ScaleImage kernels' callback (:19897-19908) reads its parameters in the Khronos order (input
// first) although the AMD kernels take the output first, so it rescales the OUTPUT's rectangle into the
// INPUT image and leaves the output's as it was.  Reproduced as such.
void scale_rect_quirk(_vx_image *in, _vx_image *out)
{
    const vx_rectangle_t r = rect_of(out);
    const float wo = (float)in->width, ho = (float)in->height, wi = (float)out->width, hi = (float)out->height;
    set_rect(in, vx_rectangle_t{(vx_uint32)(((r.start_x + 0.5f) * wo / wi) - 0.5f), (vx_uint32)(((r.start_y + 0.5f) * ho / hi) - 0.5f),
                                (vx_uint32)(((r.end_x + 0.5f) * wo / wi) - 0.5f), (vx_uint32)(((r.end_y + 0.5f) * ho / hi) - 0.5f)});
}

template <int IN, int OUT> void vr_copy(_vx_node *n) { set_rect(img(n, OUT), rect_of(img(n, IN))); }
template <int IN, int OUT, int K> void vr_shrink(_vx_node *n)
{
    if (img(n, OUT)) set_rect(img(n, OUT), shrink(rect_of(img(n, IN)), K, img(n, OUT)));
}
template <int IN, int OUT1, int OUT2, int K> void vr_shrink2(_vx_node *n)
{
    vr_shrink<IN, OUT1, K>(n);
    vr_shrink<IN, OUT2, K>(n);
}
template <int A, int B, int OUT> void vr_intersect(_vx_node *n)
{
    const vx_rectangle_t a = rect_of(img(n, A)), b = rect_of(img(n, B));
    set_rect(img(n, OUT), vx_rectangle_t{std::max(a.start_x, b.start_x), std::max(a.start_y, b.start_y), std::min(a.end_x, b.end_x),
                                         std::min(a.end_y, b.end_y)});
}

void vr_extract(_vx_node *n)
{
    _vx_image *in = img(n, 0), *o = img(n, 2);
    const vx_rectangle_t r = rect_of(in);
    if (o->width == in->width && o->height == in->height) set_rect(o, r);
    else set_rect(o, vx_rectangle_t{(r.start_x + 1) >> 1, (r.start_y + 1) >> 1, (r.end_x + 1) >> 1, (r.end_y + 1) >> 1});
}

template <int IN, int OUT> void vr_scale(_vx_node *n)
{
    _vx_image *in = img(n, IN), *o = img(n, OUT);
    if (in->width == o->width && in->height == o->height) set_rect(o, rect_of(in));   // the copy kernel
    else scale_rect_quirk(in, o);
}

void vr_pyramid(_vx_node *n)
{
    _vx_pyramid *p = (_vx_pyramid *)n->params[1];
    set_rect(p->images[0], rect_of(img(n, 0)));
    for (size_t l = 1; l < p->images.size(); l++) set_rect(p->images[l], half_rect(rect_of(p->images[l - 1]), p->images[l], false));
}

void vr_halfscale(_vx_node *n)
{
    _vx_image *in = img(n, 0), *o = img(n, 1);
    const int k = scl(n, 2) ? scl(n, 2)->value.i32 : 5;
    if (k == 1) scale_rect_quirk(in, o);
    else set_rect(o, half_rect(rect_of(in), o, k == 3));
}

// CannySobel shrinks by gradient/2, the non-max suppression by one more; the trace has no rule
void vr_canny(_vx_node *n)
{
    const int g = scl(n, 2) ? scl(n, 2)->value.i32 : 3;
    set_rect(img(n, 4), shrink(rect_of(img(n, 0)), g / 2 + 1, img(n, 4)));
}

template <int G> void vr_amd_canny_sobel(_vx_node *n) { vr_shrink<1, 0, G / 2>(n); }

ValidRectFn valid_rect_rule(vx_enum id)
{
    switch (id) {
    case VX_KERNEL_COLOR_CONVERT: return vr_copy<0, 1>;
    case VX_KERNEL_CHANNEL_EXTRACT: return vr_extract;
    case VX_KERNEL_GAUSSIAN_3x3: case VX_KERNEL_BOX_3x3: case VX_KERNEL_MEDIAN_3x3: return vr_shrink<0, 1, 1>;
    case VX_KERNEL_SOBEL_3x3: return vr_shrink2<0, 1, 2, 1>;
    case VX_KERNEL_MAGNITUDE: case VX_KERNEL_PHASE: return vr_intersect<0, 1, 2>;
    case VX_KERNEL_SCALE_IMAGE: return vr_scale<0, 1>;
    case VX_KERNEL_GAUSSIAN_PYRAMID: return vr_pyramid;
    case VX_KERNEL_HALFSCALE_GAUSSIAN: return vr_halfscale;
    case VX_KERNEL_CANNY_EDGE_DETECTOR: return vr_canny;
    case AMD_MAGNITUDE_S16_S16S16: case AMD_PHASE_U8_S16S16: return vr_intersect<1, 2, 0>;
    case AMD_CHANNEL_COPY_U8_U8: case AMD_COLOR_CONVERT_Y_RGB: return vr_copy<1, 0>;
    case AMD_BOX_U8_U8_3x3: case AMD_MEDIAN_U8_U8_3x3: case AMD_GAUSSIAN_U8_U8_3x3: return vr_shrink<1, 0, 1>;
    case AMD_SOBEL_S16S16_U8_3x3_GXY: return vr_shrink2<2, 0, 1, 1>;
    case AMD_SOBEL_MAGNITUDE_S16_U8_3x3: case AMD_SOBEL_PHASE_U8_U8_3x3: return vr_shrink<1, 0, 1>;
    case AMD_SOBEL_MAGNITUDE_PHASE_S16U8_U8_3x3: return vr_shrink2<2, 0, 1, 1>;
    case AMD_SCALE_GAUSSIAN_HALF_U8_U8_5x5: return [](_vx_node *n) { set_rect(img(n, 0), half_rect(rect_of(img(n, 1)), img(n, 0), false)); };
    case AMD_SCALE_GAUSSIAN_HALF_U8_U8_3x3: return [](_vx_node *n) { set_rect(img(n, 0), half_rect(rect_of(img(n, 1)), img(n, 0), true)); };
    case AMD_CANNY_SOBEL_U16_U8_3x3_L1NORM: case AMD_CANNY_SOBEL_U16_U8_3x3_L2NORM: return vr_amd_canny_sobel<3>;
    case AMD_CANNY_SOBEL_U16_U8_5x5_L1NORM: case AMD_CANNY_SOBEL_U16_U8_5x5_L2NORM: return vr_amd_canny_sobel<5>;
    case AMD_CANNY_SOBEL_U16_U8_7x7_L1NORM: case AMD_CANNY_SOBEL_U16_U8_7x7_L2NORM: return vr_amd_canny_sobel<7>;
    case AMD_SCALE_IMAGE_U8_U8_NEAREST: case AMD_SCALE_IMAGE_U8_U8_BILINEAR: case AMD_SCALE_IMAGE_U8_U8_BILINEAR_REPLICATE:
    case AMD_SCALE_IMAGE_U8_U8_BILINEAR_CONSTANT: case AMD_SCALE_IMAGE_U8_U8_AREA: return vr_scale<1, 0>;
    default: return nullptr;
    }
}

}  // namespace

namespace mvxe {

// called by graph verification after the node's workspace is allocated
vx_status prepare_node(_vx_node *n, hipStream_t st)
{
    if (n->kernel->id == VX_KERNEL_HARRIS_CORNERS) return i_harris_prepare(n, st);
    return VX_SUCCESS;
}

float harris_norm_factor(int gradient, int block) { return harris_norm(gradient, block); }

static std::vector<KernelDef> with_valid_rects(std::vector<KernelDef> t)
{
    for (KernelDef &k : t) k.validRect = valid_rect_rule(k.id);
    return t;
}

const std::vector<KernelDef> &kernel_table()
{
    static const std::vector<KernelDef> table = with_valid_rects({
        // Khronos vision nodes (ago_kernel_list.cpp:183-260)
        {VX_KERNEL_COLOR_CONVERT, "org.khronos.openvx.color_convert", {{IN, VX_TYPE_IMAGE, false}, {OUT, VX_TYPE_IMAGE, false}},
         v_color, nullptr, x_color, AGO_TARGET_AFFINITY_GPU},
        {VX_KERNEL_CHANNEL_EXTRACT, "org.khronos.openvx.channel_extract",
         {{IN, VX_TYPE_IMAGE, false}, {IN, VX_TYPE_SCALAR, false}, {OUT, VX_TYPE_IMAGE, false}}, v_extract, nullptr, x_extract,
         AGO_TARGET_AFFINITY_GPU},
        {VX_KERNEL_SOBEL_3x3, "org.khronos.openvx.sobel_3x3",
         {{IN, VX_TYPE_IMAGE, false}, {OUT, VX_TYPE_IMAGE, true}, {OUT, VX_TYPE_IMAGE, true}}, v_sobel, nullptr, x_sobel,
         AGO_TARGET_AFFINITY_GPU},
        {VX_KERNEL_MAGNITUDE, "org.khronos.openvx.magnitude",
         {{IN, VX_TYPE_IMAGE, false}, {IN, VX_TYPE_IMAGE, false}, {OUT, VX_TYPE_IMAGE, false}}, v_gradpair<0, VX_DF_IMAGE_S16>, nullptr,
         x_magnitude<0>, AGO_TARGET_AFFINITY_GPU},
        {VX_KERNEL_PHASE, "org.khronos.openvx.phase",
         {{IN, VX_TYPE_IMAGE, false}, {IN, VX_TYPE_IMAGE, false}, {OUT, VX_TYPE_IMAGE, false}}, v_gradpair<0, VX_DF_IMAGE_U8>, nullptr,
         x_phase<0>, AGO_TARGET_AFFINITY_GPU},
        {VX_KERNEL_SCALE_IMAGE, "org.khronos.openvx.scale_image",
         {{IN, VX_TYPE_IMAGE, false}, {OUT, VX_TYPE_IMAGE, false}, {IN, VX_TYPE_SCALAR, false}}, v_scale, nullptr, x_scale,
         AGO_TARGET_AFFINITY_GPU},
        {VX_KERNEL_MEDIAN_3x3, "org.khronos.openvx.median_3x3", {{IN, VX_TYPE_IMAGE, false}, {OUT, VX_TYPE_IMAGE, false}},
         v_filter<0>, nullptr, x_filter<0, mvx_Median_U8_U8_3x3>, AGO_TARGET_AFFINITY_GPU},
        {VX_KERNEL_BOX_3x3, "org.khronos.openvx.box_3x3", {{IN, VX_TYPE_IMAGE, false}, {OUT, VX_TYPE_IMAGE, false}}, v_filter<0>,
         nullptr, x_filter<0, mvx_Box_U8_U8_3x3>, AGO_TARGET_AFFINITY_GPU},
        {VX_KERNEL_GAUSSIAN_3x3, "org.khronos.openvx.gaussian_3x3", {{IN, VX_TYPE_IMAGE, false}, {OUT, VX_TYPE_IMAGE, false}},
         v_filter<0>, nullptr, x_filter<0, mvx_Gaussian_U8_U8_3x3>, AGO_TARGET_AFFINITY_GPU},
        {VX_KERNEL_GAUSSIAN_PYRAMID, "org.khronos.openvx.gaussian_pyramid", {{IN, VX_TYPE_IMAGE, false}, {OUT, VX_TYPE_PYRAMID, false}},
         v_pyramid, nullptr, x_pyramid, AGO_TARGET_AFFINITY_GPU},
        {VX_KERNEL_CANNY_EDGE_DETECTOR, "org.khronos.openvx.canny_edge_detector",
         {{IN, VX_TYPE_IMAGE, false}, {IN, VX_TYPE_THRESHOLD, false}, {IN, VX_TYPE_SCALAR, false}, {IN, VX_TYPE_SCALAR, false},
          {OUT, VX_TYPE_IMAGE, false}},
         v_canny, i_canny, x_canny, AGO_TARGET_AFFINITY_GPU},
        {VX_KERNEL_HARRIS_CORNERS, "org.khronos.openvx.harris_corners",
         {{IN, VX_TYPE_IMAGE, false}, {IN, VX_TYPE_SCALAR, false}, {IN, VX_TYPE_SCALAR, false}, {IN, VX_TYPE_SCALAR, false},
          {IN, VX_TYPE_SCALAR, false}, {IN, VX_TYPE_SCALAR, false}, {OUT, VX_TYPE_ARRAY, false}, {OUT, VX_TYPE_SCALAR, true}},
         v_harris, i_harris, x_harris, AGO_TARGET_AFFINITY_GPU},
        {VX_KERNEL_FAST_CORNERS, "org.khronos.openvx.fast_corners",
         {{IN, VX_TYPE_IMAGE, false}, {IN, VX_TYPE_SCALAR, false}, {IN, VX_TYPE_SCALAR, false}, {OUT, VX_TYPE_ARRAY, false},
          {OUT, VX_TYPE_SCALAR, true}},
         v_fast, i_fast, x_fast, AGO_TARGET_AFFINITY_GPU},
        {AMD_FAST_CORNERS_XY_U8_SUPRESSION, "com.amd.openvx.FastCorners_XY_U8_Supression",
         {{OUT, VX_TYPE_ARRAY, false}, {OUT, VX_TYPE_SCALAR, true}, {IN, VX_TYPE_IMAGE, false}, {IN, VX_TYPE_SCALAR, false}},
         v_amd_fast, i_amd_fast, x_amd_fast<true>, AGO_TARGET_AFFINITY_GPU},
        {AMD_FAST_CORNERS_XY_U8_NOSUPRESSION, "com.amd.openvx.FastCorners_XY_U8_NoSupression",
         {{OUT, VX_TYPE_ARRAY, false}, {OUT, VX_TYPE_SCALAR, true}, {IN, VX_TYPE_IMAGE, false}, {IN, VX_TYPE_SCALAR, false}},
         v_amd_fast, i_amd_fast, x_amd_fast<false>, AGO_TARGET_AFFINITY_GPU},
        {VX_KERNEL_WARP_AFFINE, "org.khronos.openvx.warp_affine",
         {{IN, VX_TYPE_IMAGE, false}, {IN, VX_TYPE_MATRIX, false}, {IN, VX_TYPE_SCALAR, false}, {OUT, VX_TYPE_IMAGE, false}},
         v_warp<false>, nullptr, x_warp<false>, AGO_TARGET_AFFINITY_GPU},
        {VX_KERNEL_WARP_PERSPECTIVE, "org.khronos.openvx.warp_perspective",
         {{IN, VX_TYPE_IMAGE, false}, {IN, VX_TYPE_MATRIX, false}, {IN, VX_TYPE_SCALAR, false}, {OUT, VX_TYPE_IMAGE, false}},
         v_warp<true>, nullptr, x_warp<true>, AGO_TARGET_AFFINITY_GPU},
        {VX_KERNEL_OPTICAL_FLOW_PYR_LK, "org.khronos.openvx.optical_flow_pyr_lk",
         {{IN, VX_TYPE_PYRAMID, false}, {IN, VX_TYPE_PYRAMID, false}, {IN, VX_TYPE_ARRAY, false}, {IN, VX_TYPE_ARRAY, false},
          {OUT, VX_TYPE_ARRAY, false}, {IN, VX_TYPE_SCALAR, false}, {IN, VX_TYPE_SCALAR, false}, {IN, VX_TYPE_SCALAR, false},
          {IN, VX_TYPE_SCALAR, false}, {IN, VX_TYPE_SCALAR, false}},
         v_lk, i_lk, x_lk, AGO_TARGET_AFFINITY_GPU},
        {VX_KERNEL_HALFSCALE_GAUSSIAN, "org.khronos.openvx.halfscale_gaussian",
         {{IN, VX_TYPE_IMAGE, false}, {OUT, VX_TYPE_IMAGE, false}, {IN, VX_TYPE_SCALAR, false}}, v_halfscale, nullptr, x_halfscale,
         AGO_TARGET_AFFINITY_GPU},
        {VX_KERNEL_REMAP, "org.khronos.openvx.remap",
         {{IN, VX_TYPE_IMAGE, false}, {IN, VX_TYPE_REMAP, false}, {IN, VX_TYPE_SCALAR, false}, {OUT, VX_TYPE_IMAGE, false}},
         v_remap, nullptr, x_remap, AGO_TARGET_AFFINITY_GPU},
        // AMD low-level kernels, output first (ago_kernel_list.cpp:339-474)
        {AMD_WARP_AFFINE_U8_U8_NEAREST, "com.amd.openvx.WarpAffine_U8_U8_Nearest",
         {{OUT, VX_TYPE_IMAGE, false}, {IN, VX_TYPE_IMAGE, false}, {IN, VX_TYPE_MATRIX, false}}, v_amd_warp<false, false>, nullptr,
         x_amd_warp<false, false, true>, AGO_TARGET_AFFINITY_GPU},
        {AMD_WARP_AFFINE_U8_U8_NEAREST_CONSTANT, "com.amd.openvx.WarpAffine_U8_U8_Nearest_Constant",
         {{OUT, VX_TYPE_IMAGE, false}, {IN, VX_TYPE_IMAGE, false}, {IN, VX_TYPE_MATRIX, false}, {IN, VX_TYPE_SCALAR, false}},
         v_amd_warp<false, true>, nullptr, x_amd_warp<false, true, true>, AGO_TARGET_AFFINITY_GPU},
        {AMD_WARP_AFFINE_U8_U8_BILINEAR, "com.amd.openvx.WarpAffine_U8_U8_Bilinear",
         {{OUT, VX_TYPE_IMAGE, false}, {IN, VX_TYPE_IMAGE, false}, {IN, VX_TYPE_MATRIX, false}}, v_amd_warp<false, false>, nullptr,
         x_amd_warp<false, false, false>, AGO_TARGET_AFFINITY_GPU},
        {AMD_WARP_AFFINE_U8_U8_BILINEAR_CONSTANT, "com.amd.openvx.WarpAffine_U8_U8_Bilinear_Constant",
         {{OUT, VX_TYPE_IMAGE, false}, {IN, VX_TYPE_IMAGE, false}, {IN, VX_TYPE_MATRIX, false}, {IN, VX_TYPE_SCALAR, false}},
         v_amd_warp<false, true>, nullptr, x_amd_warp<false, true, false>, AGO_TARGET_AFFINITY_GPU},
        {AMD_WARP_PERSPECTIVE_U8_U8_NEAREST, "com.amd.openvx.WarpPerspective_U8_U8_Nearest",
         {{OUT, VX_TYPE_IMAGE, false}, {IN, VX_TYPE_IMAGE, false}, {IN, VX_TYPE_MATRIX, false}}, v_amd_warp<true, false>, nullptr,
         x_amd_warp<true, false, true>, AGO_TARGET_AFFINITY_GPU},
        {AMD_WARP_PERSPECTIVE_U8_U8_NEAREST_CONSTANT, "com.amd.openvx.WarpPerspective_U8_U8_Nearest_Constant",
         {{OUT, VX_TYPE_IMAGE, false}, {IN, VX_TYPE_IMAGE, false}, {IN, VX_TYPE_MATRIX, false}, {IN, VX_TYPE_SCALAR, false}},
         v_amd_warp<true, true>, nullptr, x_amd_warp<true, true, true>, AGO_TARGET_AFFINITY_GPU},
        {AMD_WARP_PERSPECTIVE_U8_U8_BILINEAR, "com.amd.openvx.WarpPerspective_U8_U8_Bilinear",
         {{OUT, VX_TYPE_IMAGE, false}, {IN, VX_TYPE_IMAGE, false}, {IN, VX_TYPE_MATRIX, false}}, v_amd_warp<true, false>, nullptr,
         x_amd_warp<true, false, false>, AGO_TARGET_AFFINITY_GPU},
        {AMD_WARP_PERSPECTIVE_U8_U8_BILINEAR_CONSTANT, "com.amd.openvx.WarpPerspective_U8_U8_Bilinear_Constant",
         {{OUT, VX_TYPE_IMAGE, false}, {IN, VX_TYPE_IMAGE, false}, {IN, VX_TYPE_MATRIX, false}, {IN, VX_TYPE_SCALAR, false}},
         v_amd_warp<true, true>, nullptr, x_amd_warp<true, true, false>, AGO_TARGET_AFFINITY_GPU},
        {AMD_REMAP_U8_U8_NEAREST, "com.amd.openvx.Remap_U8_U8_Nearest",
         {{OUT, VX_TYPE_IMAGE, false}, {IN, VX_TYPE_IMAGE, false}, {IN, VX_TYPE_REMAP, false}}, v_amd_remap<false>, nullptr,
         x_amd_remap<false, true>, AGO_TARGET_AFFINITY_GPU},
        {AMD_REMAP_U8_U8_NEAREST_CONSTANT, "com.amd.openvx.Remap_U8_U8_Nearest_Constant",
         {{OUT, VX_TYPE_IMAGE, false}, {IN, VX_TYPE_IMAGE, false}, {IN, VX_TYPE_REMAP, false}, {IN, VX_TYPE_SCALAR, false}},
         v_amd_remap<true>, nullptr, x_amd_remap<true, true>, AGO_TARGET_AFFINITY_GPU},
        {AMD_REMAP_U8_U8_BILINEAR, "com.amd.openvx.Remap_U8_U8_Bilinear",
         {{OUT, VX_TYPE_IMAGE, false}, {IN, VX_TYPE_IMAGE, false}, {IN, VX_TYPE_REMAP, false}}, v_amd_remap<false>, nullptr,
         x_amd_remap<false, false>, AGO_TARGET_AFFINITY_GPU},
        {AMD_REMAP_U8_U8_BILINEAR_CONSTANT, "com.amd.openvx.Remap_U8_U8_Bilinear_Constant",
         {{OUT, VX_TYPE_IMAGE, false}, {IN, VX_TYPE_IMAGE, false}, {IN, VX_TYPE_REMAP, false}, {IN, VX_TYPE_SCALAR, false}},
         v_amd_remap<true>, nullptr, x_amd_remap<true, false>, AGO_TARGET_AFFINITY_GPU},
        {AMD_MAGNITUDE_S16_S16S16, "com.amd.openvx.Magnitude_S16_S16S16",
         {{OUT, VX_TYPE_IMAGE, false}, {IN, VX_TYPE_IMAGE, false}, {IN, VX_TYPE_IMAGE, false}}, v_gradpair<1, VX_DF_IMAGE_S16>, nullptr,
         x_magnitude<1>, AGO_TARGET_AFFINITY_GPU},
        {AMD_PHASE_U8_S16S16, "com.amd.openvx.Phase_U8_S16S16",
         {{OUT, VX_TYPE_IMAGE, false}, {IN, VX_TYPE_IMAGE, false}, {IN, VX_TYPE_IMAGE, false}}, v_gradpair<1, VX_DF_IMAGE_U8>, nullptr,
         x_phase<1>, AGO_TARGET_AFFINITY_GPU},
        {AMD_CHANNEL_COPY_U8_U8, "com.amd.openvx.ChannelCopy_U8_U8", {{OUT, VX_TYPE_IMAGE, false}, {IN, VX_TYPE_IMAGE, false}},
         v_amd_copy, nullptr, x_amd_copy, AGO_TARGET_AFFINITY_GPU},
        {AMD_COLOR_CONVERT_Y_RGB, "com.amd.openvx.ColorConvert_Y_RGB", {{OUT, VX_TYPE_IMAGE, false}, {IN, VX_TYPE_IMAGE, false}},
         v_amd_color_y, nullptr, x_amd_color_y, AGO_TARGET_AFFINITY_GPU},
        {AMD_BOX_U8_U8_3x3, "com.amd.openvx.Box_U8_U8_3x3", {{OUT, VX_TYPE_IMAGE, false}, {IN, VX_TYPE_IMAGE, false}}, v_filter<1>,
         nullptr, x_filter<1, mvx_Box_U8_U8_3x3>, AGO_TARGET_AFFINITY_GPU},
        {AMD_MEDIAN_U8_U8_3x3, "com.amd.openvx.Median_U8_U8_3x3", {{OUT, VX_TYPE_IMAGE, false}, {IN, VX_TYPE_IMAGE, false}},
         v_filter<1>, nullptr, x_filter<1, mvx_Median_U8_U8_3x3>, AGO_TARGET_AFFINITY_GPU},
        {AMD_GAUSSIAN_U8_U8_3x3, "com.amd.openvx.Gaussian_U8_U8_3x3", {{OUT, VX_TYPE_IMAGE, false}, {IN, VX_TYPE_IMAGE, false}},
         v_filter<1>, nullptr, x_filter<1, mvx_Gaussian_U8_U8_3x3>, AGO_TARGET_AFFINITY_GPU},
        {AMD_SCALE_GAUSSIAN_HALF_U8_U8_3x3, "com.amd.openvx.ScaleGaussianHalf_U8_U8_3x3",
         {{OUT, VX_TYPE_IMAGE, false}, {IN, VX_TYPE_IMAGE, false}}, v_amd_half<3>, nullptr, x_amd_half<3>, AGO_TARGET_AFFINITY_GPU},
        {AMD_SCALE_GAUSSIAN_HALF_U8_U8_5x5, "com.amd.openvx.ScaleGaussianHalf_U8_U8_5x5",
         {{OUT, VX_TYPE_IMAGE, false}, {IN, VX_TYPE_IMAGE, false}}, v_amd_half<5>, nullptr, x_amd_half<5>, AGO_TARGET_AFFINITY_GPU},
        {AMD_SOBEL_S16S16_U8_3x3_GXY, "com.amd.openvx.Sobel_S16S16_U8_3x3_GXY",
         {{OUT, VX_TYPE_IMAGE, false}, {OUT, VX_TYPE_IMAGE, false}, {IN, VX_TYPE_IMAGE, false}}, v_amd_sobel, nullptr, x_amd_sobel,
         AGO_TARGET_AFFINITY_GPU},
        {AMD_SOBEL_MAGNITUDE_S16_U8_3x3, "com.amd.openvx.SobelMagnitude_S16_U8_3x3",
         {{OUT, VX_TYPE_IMAGE, false}, {IN, VX_TYPE_IMAGE, false}}, v_amd_sobel_merged<1>, nullptr,
         x_amd_sobel_merged<AMD_SOBEL_MAGNITUDE_S16_U8_3x3>, AGO_TARGET_AFFINITY_GPU},
        {AMD_SOBEL_PHASE_U8_U8_3x3, "com.amd.openvx.SobelPhase_U8_U8_3x3",
         {{OUT, VX_TYPE_IMAGE, false}, {IN, VX_TYPE_IMAGE, false}}, v_amd_sobel_merged<1>, nullptr,
         x_amd_sobel_merged<AMD_SOBEL_PHASE_U8_U8_3x3>, AGO_TARGET_AFFINITY_GPU},
        {AMD_SOBEL_MAGNITUDE_PHASE_S16U8_U8_3x3, "com.amd.openvx.SobelMagnitudePhase_S16U8_U8_3x3",
         {{OUT, VX_TYPE_IMAGE, false}, {OUT, VX_TYPE_IMAGE, false}, {IN, VX_TYPE_IMAGE, false}}, v_amd_sobel_merged<2>, nullptr,
         x_amd_sobel_merged<AMD_SOBEL_MAGNITUDE_PHASE_S16U8_U8_3x3>, AGO_TARGET_AFFINITY_GPU},
        {AMD_CANNY_SOBEL_U16_U8_3x3_L1NORM, "com.amd.openvx.CannySobel_U16_U8_3x3_L1NORM",
         {{OUT, VX_TYPE_IMAGE, false}, {IN, VX_TYPE_IMAGE, false}}, v_amd_canny_sobel, nullptr, x_amd_canny_sobel<3, false>,
         AGO_TARGET_AFFINITY_GPU},
        {AMD_CANNY_SOBEL_U16_U8_3x3_L2NORM, "com.amd.openvx.CannySobel_U16_U8_3x3_L2NORM",
         {{OUT, VX_TYPE_IMAGE, false}, {IN, VX_TYPE_IMAGE, false}}, v_amd_canny_sobel, nullptr, x_amd_canny_sobel<3, true>,
         AGO_TARGET_AFFINITY_GPU},
        {AMD_CANNY_SOBEL_U16_U8_5x5_L1NORM, "com.amd.openvx.CannySobel_U16_U8_5x5_L1NORM",
         {{OUT, VX_TYPE_IMAGE, false}, {IN, VX_TYPE_IMAGE, false}}, v_amd_canny_sobel, nullptr, x_amd_canny_sobel<5, false>,
         AGO_TARGET_AFFINITY_GPU},
        {AMD_CANNY_SOBEL_U16_U8_5x5_L2NORM, "com.amd.openvx.CannySobel_U16_U8_5x5_L2NORM",
         {{OUT, VX_TYPE_IMAGE, false}, {IN, VX_TYPE_IMAGE, false}}, v_amd_canny_sobel, nullptr, x_amd_canny_sobel<5, true>,
         AGO_TARGET_AFFINITY_GPU},
        {AMD_CANNY_SOBEL_U16_U8_7x7_L1NORM, "com.amd.openvx.CannySobel_U16_U8_7x7_L1NORM",
         {{OUT, VX_TYPE_IMAGE, false}, {IN, VX_TYPE_IMAGE, false}}, v_amd_canny_sobel, nullptr, x_amd_canny_sobel<7, false>,
         AGO_TARGET_AFFINITY_GPU},
        {AMD_CANNY_SOBEL_U16_U8_7x7_L2NORM, "com.amd.openvx.CannySobel_U16_U8_7x7_L2NORM",
         {{OUT, VX_TYPE_IMAGE, false}, {IN, VX_TYPE_IMAGE, false}}, v_amd_canny_sobel, nullptr, x_amd_canny_sobel<7, true>,
         AGO_TARGET_AFFINITY_GPU},
        {AMD_SCALE_IMAGE_U8_U8_NEAREST, "com.amd.openvx.ScaleImage_U8_U8_Nearest",
         {{OUT, VX_TYPE_IMAGE, false}, {IN, VX_TYPE_IMAGE, false}}, v_amd_scale_k<0>, nullptr, x_amd_scale<0>, AGO_TARGET_AFFINITY_GPU},
        {AMD_SCALE_IMAGE_U8_U8_BILINEAR, "com.amd.openvx.ScaleImage_U8_U8_Bilinear",
         {{OUT, VX_TYPE_IMAGE, false}, {IN, VX_TYPE_IMAGE, false}}, v_amd_scale_k<1>, nullptr, x_amd_scale<1>, AGO_TARGET_AFFINITY_GPU},
        {AMD_SCALE_IMAGE_U8_U8_BILINEAR_REPLICATE, "com.amd.openvx.ScaleImage_U8_U8_Bilinear_Replicate",
         {{OUT, VX_TYPE_IMAGE, false}, {IN, VX_TYPE_IMAGE, false}}, v_amd_scale_k<2>, nullptr, x_amd_scale<2>, AGO_TARGET_AFFINITY_GPU},
        {AMD_SCALE_IMAGE_U8_U8_BILINEAR_CONSTANT, "com.amd.openvx.ScaleImage_U8_U8_Bilinear_Constant",
         {{OUT, VX_TYPE_IMAGE, false}, {IN, VX_TYPE_IMAGE, false}, {IN, VX_TYPE_SCALAR, false}}, v_amd_scale_k<3>, nullptr,
         x_amd_scale<3>, AGO_TARGET_AFFINITY_GPU},
        {AMD_SCALE_IMAGE_U8_U8_AREA, "com.amd.openvx.ScaleImage_U8_U8_Area",
         {{OUT, VX_TYPE_IMAGE, false}, {IN, VX_TYPE_IMAGE, false}}, v_amd_scale_k<4>, nullptr, x_amd_scale<4>, AGO_TARGET_AFFINITY_GPU},
    });
    return table;
}

}  // namespace mvxe
