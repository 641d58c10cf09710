// Test-infrastructure harness around the reference's own CPU engine (oracle/_ref/libopenvx_ref.so,
// compiled from the reference sources by oracle/Makefile).  Each entry point builds a small OpenVX
// graph with the reference's public API, runs it on the reference CPU path (AGO + ago_haf_cpu*),
// and copies every output plane out to a tightly packed host buffer (stride = width * bytes/pixel).
//
// Output images are created fresh by each call, so pixels a reference kernel never writes
// (rows 0 and h-1 of a 3x3 filter, for example) come back as the zeros of the reference's
// calloc'd buffers (ago_util.cpp:187-200).
//
// Nothing in the product links this file; tests and the CPU-baseline leg of bench.py load it
// through ctypes when it has been built.
#include <VX/vx.h>
#include <VX/vxu.h>
#include <vx_ext_amd.h>
#include <cstdint>
#include <cstring>
#include <cstdio>
#include <cmath>
#include <vector>
#include <chrono>

namespace {

struct Ctx {
    vx_context ctx;
    vx_graph graph;
    Ctx() { ctx = vxCreateContext(); graph = vxCreateGraph(ctx); }
    ~Ctx() { vxReleaseGraph(&graph); vxReleaseContext(&ctx); }
};

int bpp_of(vx_df_image f) {
    switch (f) {
    case VX_DF_IMAGE_U8: return 1;
    case VX_DF_IMAGE_S16: case VX_DF_IMAGE_U16: return 2;
    case VX_DF_IMAGE_RGB: return 3;
    case VX_DF_IMAGE_RGBX: case VX_DF_IMAGE_U32: case VX_DF_IMAGE_S32: return 4;
    default: return 1;
    }
}

vx_status put(vx_image img, int w, int h, const void *src, int bpp) {
    vx_rectangle_t r = {0, 0, (vx_uint32)w, (vx_uint32)h};
    vx_imagepatch_addressing_t a = {};
    a.dim_x = w; a.dim_y = h; a.stride_x = bpp; a.stride_y = w * bpp;
    a.scale_x = VX_SCALE_UNITY; a.scale_y = VX_SCALE_UNITY; a.step_x = 1; a.step_y = 1;
    return vxCopyImagePatch(img, &r, 0, &a, (void *)src, VX_WRITE_ONLY, VX_MEMORY_TYPE_HOST);
}

vx_status get(vx_image img, int w, int h, void *dst, int bpp, int plane = 0) {
    vx_rectangle_t r = {0, 0, (vx_uint32)w, (vx_uint32)h};
    vx_imagepatch_addressing_t a = {};
    a.dim_x = w; a.dim_y = h; a.stride_x = bpp; a.stride_y = w * bpp;
    a.scale_x = VX_SCALE_UNITY; a.scale_y = VX_SCALE_UNITY; a.step_x = 1; a.step_y = 1;
    return vxCopyImagePatch(img, &r, plane, &a, dst, VX_READ_ONLY, VX_MEMORY_TYPE_HOST);
}

// U8 rows including their padding: the image's own buffer (vxMapImagePatch returns a pointer into it)
// read over ALIGN16(w) bytes per row, so the bytes the reference's 4-pixel stores leave past the
// width are seen too
vx_status get_padded(vx_image img, int w, int h, uint8_t *dst) {
    vx_rectangle_t r = {0, 0, (vx_uint32)w, (vx_uint32)h};
    vx_imagepatch_addressing_t a = {};
    vx_map_id id;
    void *ptr = nullptr;
    vx_status s = vxMapImagePatch(img, &r, 0, &id, &a, &ptr, VX_READ_ONLY, VX_MEMORY_TYPE_HOST, 0);
    if (s != VX_SUCCESS) return s;
    const int ds = (w + 15) & ~15;
    if (a.stride_y < ds) { vxUnmapImagePatch(img, id); return VX_FAILURE; }
    for (int y = 0; y < h; y++) memcpy(dst + (size_t)y * ds, (const uint8_t *)ptr + (size_t)y * a.stride_y, (size_t)ds);
    return vxUnmapImagePatch(img, id);
}

vx_image make_input(Ctx &c, int w, int h, vx_df_image f, const void *src) {
    vx_image img = vxCreateImage(c.ctx, w, h, f);
    if (put(img, w, h, src, bpp_of(f)) != VX_SUCCESS) return nullptr;
    return img;
}

#define CHECK(x) do { vx_status _s = (x); if (_s != VX_SUCCESS) { fprintf(stderr, "ref_harness: %s -> %d (line %d)\n", #x, _s, __LINE__); return _s; } } while (0)

int run(Ctx &c) {
    CHECK(vxVerifyGraph(c.graph));
    CHECK(vxProcessGraph(c.graph));
    return 0;
}

}  // namespace

extern "C" {

// U8 -> U8 3x3 filters (vx_nodes.cpp vxGaussian3x3Node/vxBox3x3Node/vxMedian3x3Node).
// which: 0 gaussian, 1 box, 2 median.
int ref_filter3x3(int which, int w, int h, const uint8_t *src, uint8_t *dst) {
    Ctx c;
    vx_image in = make_input(c, w, h, VX_DF_IMAGE_U8, src);
    vx_image out = vxCreateImage(c.ctx, w, h, VX_DF_IMAGE_U8);
    vx_node n = which == 0 ? vxGaussian3x3Node(c.graph, in, out)
              : which == 1 ? vxBox3x3Node(c.graph, in, out)
                           : vxMedian3x3Node(c.graph, in, out);
    if (!n) return -1;
    int s = run(c); if (s) return s;
    CHECK(get(out, w, h, dst, 1));
    vxReleaseNode(&n); vxReleaseImage(&in); vxReleaseImage(&out);
    return 0;
}

// Sobel3x3: U8 -> S16 gx, gy (either may be null to leave that output unconnected).
int ref_sobel3x3(int w, int h, const uint8_t *src, int16_t *gx, int16_t *gy) {
    Ctx c;
    vx_image in = make_input(c, w, h, VX_DF_IMAGE_U8, src);
    vx_image ox = gx ? vxCreateImage(c.ctx, w, h, VX_DF_IMAGE_S16) : nullptr;
    vx_image oy = gy ? vxCreateImage(c.ctx, w, h, VX_DF_IMAGE_S16) : nullptr;
    vx_node n = vxSobel3x3Node(c.graph, in, ox, oy);
    if (!n) return -1;
    int s = run(c); if (s) return s;
    if (gx) CHECK(get(ox, w, h, gx, 2));
    if (gy) CHECK(get(oy, w, h, gy, 2));
    return 0;
}

// Magnitude (S16, S16 -> S16) and Phase (S16, S16 -> U8); mag or phase may be null.
int ref_magnitude_phase(int w, int h, const int16_t *gx, const int16_t *gy, int16_t *mag, uint8_t *phase) {
    Ctx c;
    vx_image ix = make_input(c, w, h, VX_DF_IMAGE_S16, gx);
    vx_image iy = make_input(c, w, h, VX_DF_IMAGE_S16, gy);
    vx_image om = mag ? vxCreateImage(c.ctx, w, h, VX_DF_IMAGE_S16) : nullptr;
    vx_image op = phase ? vxCreateImage(c.ctx, w, h, VX_DF_IMAGE_U8) : nullptr;
    if (mag && !vxMagnitudeNode(c.graph, ix, iy, om)) return -1;
    if (phase && !vxPhaseNode(c.graph, ix, iy, op)) return -1;
    int s = run(c); if (s) return s;
    if (mag) CHECK(get(om, w, h, mag, 2));
    if (phase) CHECK(get(op, w, h, phase, 1));
    return 0;
}

// The plain CPU-runnable graph: Gaussian3x3 (virtual) -> Sobel3x3 (gx, gy) -> Magnitude + Phase,
// border UNDEFINED.  If time_iters > 0 the graph is processed that many more times and the
// mean seconds per vxProcessGraph is stored in *sec_per_frame.
int ref_chain1(int w, int h, const uint8_t *src, int16_t *gx, int16_t *gy, int16_t *mag, uint8_t *phase,
               int time_iters, double *sec_per_frame) {
    Ctx c;
    vx_image in = make_input(c, w, h, VX_DF_IMAGE_U8, src);
    vx_image g = vxCreateVirtualImage(c.graph, w, h, VX_DF_IMAGE_U8);
    vx_image ox = vxCreateImage(c.ctx, w, h, VX_DF_IMAGE_S16);
    vx_image oy = vxCreateImage(c.ctx, w, h, VX_DF_IMAGE_S16);
    vx_image om = vxCreateImage(c.ctx, w, h, VX_DF_IMAGE_S16);
    vx_image op = vxCreateImage(c.ctx, w, h, VX_DF_IMAGE_U8);
    if (!vxGaussian3x3Node(c.graph, in, g)) return -1;
    if (!vxSobel3x3Node(c.graph, g, ox, oy)) return -1;
    if (!vxMagnitudeNode(c.graph, ox, oy, om)) return -1;
    if (!vxPhaseNode(c.graph, ox, oy, op)) return -1;
    int s = run(c); if (s) return s;
    CHECK(get(ox, w, h, gx, 2));
    CHECK(get(oy, w, h, gy, 2));
    CHECK(get(om, w, h, mag, 2));
    CHECK(get(op, w, h, phase, 1));
    if (time_iters > 0 && sec_per_frame) {
        auto t0 = std::chrono::steady_clock::now();
        for (int i = 0; i < time_iters; i++) CHECK(vxProcessGraph(c.graph));
        auto t1 = std::chrono::steady_clock::now();
        *sec_per_frame = std::chrono::duration<double>(t1 - t0).count() / time_iters;
    }
    return 0;
}

// [Gaussian3x3 (virtual) ->] Sobel3x3 with VIRTUAL gx and gy -> Magnitude and/or Phase, so that the
// reference's merge rules (ago_drama_remove.cpp:155-182) pick SobelMagnitudePhase / SobelMagnitude /
// SobelPhase.  which: 0 magnitude + phase, 1 magnitude only, 2 phase only.
int ref_sobel_merged(int which, int blur, int w, int h, const uint8_t *src, int16_t *mag, uint8_t *phase) {
    Ctx c;
    vx_image in = make_input(c, w, h, VX_DF_IMAGE_U8, src);
    vx_image s = in;
    if (blur) {
        s = vxCreateVirtualImage(c.graph, w, h, VX_DF_IMAGE_U8);
        if (!vxGaussian3x3Node(c.graph, in, s)) return -1;
    }
    vx_image gx = vxCreateVirtualImage(c.graph, w, h, VX_DF_IMAGE_S16);
    vx_image gy = vxCreateVirtualImage(c.graph, w, h, VX_DF_IMAGE_S16);
    vx_image om = which != 2 ? vxCreateImage(c.ctx, w, h, VX_DF_IMAGE_S16) : nullptr;
    vx_image op = which != 1 ? vxCreateImage(c.ctx, w, h, VX_DF_IMAGE_U8) : nullptr;
    if (!vxSobel3x3Node(c.graph, s, gx, gy)) return -1;
    if (om && !vxMagnitudeNode(c.graph, gx, gy, om)) return -1;
    if (op && !vxPhaseNode(c.graph, gx, gy, op)) return -1;
    int st = run(c); if (st) return st;
    if (om) CHECK(get(om, w, h, mag, 2));
    if (op) CHECK(get(op, w, h, phase, 1));
    return 0;
}

// The uneven-batch chain on one image: Gaussian3x3 (virtual) -> Sobel3x3 -> Magnitude + Phase and
// ScaleImage bilinear of the input to (w/2, h/2), one graph; timing as ref_chain1.
int ref_chain_half(int w, int h, const uint8_t *src, int16_t *gx, int16_t *gy, int16_t *mag, uint8_t *phase, uint8_t *half,
                   int time_iters, double *sec_per_frame) {
    Ctx c;
    vx_image in = make_input(c, w, h, VX_DF_IMAGE_U8, src);
    vx_image g = vxCreateVirtualImage(c.graph, w, h, VX_DF_IMAGE_U8);
    vx_image ox = vxCreateImage(c.ctx, w, h, VX_DF_IMAGE_S16);
    vx_image oy = vxCreateImage(c.ctx, w, h, VX_DF_IMAGE_S16);
    vx_image om = vxCreateImage(c.ctx, w, h, VX_DF_IMAGE_S16);
    vx_image op = vxCreateImage(c.ctx, w, h, VX_DF_IMAGE_U8);
    vx_image oh = vxCreateImage(c.ctx, w / 2, h / 2, VX_DF_IMAGE_U8);
    if (!vxGaussian3x3Node(c.graph, in, g)) return -1;
    if (!vxSobel3x3Node(c.graph, g, ox, oy)) return -1;
    if (!vxMagnitudeNode(c.graph, ox, oy, om)) return -1;
    if (!vxPhaseNode(c.graph, ox, oy, op)) return -1;
    if (!vxScaleImageNode(c.graph, in, oh, VX_INTERPOLATION_BILINEAR)) return -1;
    int s = run(c); if (s) return s;
    CHECK(get(ox, w, h, gx, 2));
    CHECK(get(oy, w, h, gy, 2));
    CHECK(get(om, w, h, mag, 2));
    CHECK(get(op, w, h, phase, 1));
    CHECK(get(oh, w / 2, h / 2, half, 1));
    if (time_iters > 0 && sec_per_frame) {
        auto t0 = std::chrono::steady_clock::now();
        for (int i = 0; i < time_iters; i++) CHECK(vxProcessGraph(c.graph));
        auto t1 = std::chrono::steady_clock::now();
        *sec_per_frame = std::chrono::duration<double>(t1 - t0).count() / time_iters;
    }
    return 0;
}

// ColorConvert RGB -> IYUV (virtual) -> ChannelExtract Y.
int ref_color_y(int w, int h, const uint8_t *rgb, uint8_t *y) {
    Ctx c;
    vx_image in = make_input(c, w, h, VX_DF_IMAGE_RGB, rgb);
    vx_image iyuv = vxCreateVirtualImage(c.graph, w, h, VX_DF_IMAGE_IYUV);
    vx_image oy = vxCreateImage(c.ctx, w, h, VX_DF_IMAGE_U8);
    if (!vxColorConvertNode(c.graph, in, iyuv)) return -1;
    if (!vxChannelExtractNode(c.graph, iyuv, VX_CHANNEL_Y, oy)) return -1;
    int s = run(c); if (s) return s;
    CHECK(get(oy, w, h, y, 1));
    return 0;
}

// ScaleImage bilinear (border UNDEFINED).
int ref_scale(int sw, int sh, const uint8_t *src, int dw, int dh, uint8_t *dst) {
    Ctx c;
    vx_image in = make_input(c, sw, sh, VX_DF_IMAGE_U8, src);
    vx_image out = vxCreateImage(c.ctx, dw, dh, VX_DF_IMAGE_U8);
    if (!vxScaleImageNode(c.graph, in, out, VX_INTERPOLATION_BILINEAR)) return -1;
    int s = run(c); if (s) return s;
    CHECK(get(out, dw, dh, dst, 1));
    return 0;
}

// ScaleImage with any interpolation (VX_INTERPOLATION_*) and node border mode (VX_BORDER_*; the
// constant value for VX_BORDER_CONSTANT).  The output image is fresh (zeros where nothing writes).
int ref_scale_mode(int sw, int sh, const uint8_t *src, int dw, int dh, uint8_t *dst, int interp, int border_mode,
                   int border_value) {
    Ctx c;
    vx_image in = make_input(c, sw, sh, VX_DF_IMAGE_U8, src);
    vx_image out = vxCreateImage(c.ctx, dw, dh, VX_DF_IMAGE_U8);
    vx_node n = vxScaleImageNode(c.graph, in, out, (vx_enum)interp);
    if (!n) return -1;
    vx_border_t b = {};
    b.mode = (vx_enum)border_mode;
    b.constant_value.U8 = (vx_uint8)border_value;
    CHECK(vxSetNodeAttribute(n, VX_NODE_BORDER, &b, sizeof(b)));
    int s = run(c); if (s) return s;
    CHECK(get(out, dw, dh, dst, 1));
    return 0;
}

// The AMD kernel com.amd.openvx.ScaleImage_U8_U8_Area run by name (parameters: output, input), which
// reaches its same-size form -- vxScaleImageNode turns a same-size scale into a plane copy instead.
int ref_amd_scale_area(int sw, int sh, const uint8_t *src, int dw, int dh, uint8_t *dst) {
    Ctx c;
    vx_image in = make_input(c, sw, sh, VX_DF_IMAGE_U8, src);
    vx_image out = vxCreateImage(c.ctx, dw, dh, VX_DF_IMAGE_U8);
    vx_kernel k = vxGetKernelByName(c.ctx, "com.amd.openvx.ScaleImage_U8_U8_Area");
    if (!k) return -1;
    vx_node n = vxCreateGenericNode(c.graph, k);
    if (!n) return -1;
    CHECK(vxSetParameterByIndex(n, 0, (vx_reference)out));
    CHECK(vxSetParameterByIndex(n, 1, (vx_reference)in));
    int s = run(c); if (s) return s;
    CHECK(get(out, dw, dh, dst, 1));
    return 0;
}

// color_convert IYUV -> RGB (nv12 == 0: y, u, v planes of w x h, w/2 x h/2, w/2 x h/2) or NV12 -> RGB
// (nv12 != 0: y plane and the interleaved w/2 x h/2 UV plane passed as u).
int ref_color_rgb_yuv(int w, int h, const uint8_t *y, const uint8_t *u, const uint8_t *v, int nv12, uint8_t *rgb) {
    Ctx c;
    vx_image in = vxCreateImage(c.ctx, w, h, nv12 ? VX_DF_IMAGE_NV12 : VX_DF_IMAGE_IYUV);
    vx_rectangle_t r = {0, 0, (vx_uint32)w, (vx_uint32)h};
    vx_imagepatch_addressing_t a = {};
    a.dim_x = w; a.dim_y = h; a.stride_x = 1; a.stride_y = w;
    a.scale_x = VX_SCALE_UNITY; a.scale_y = VX_SCALE_UNITY; a.step_x = 1; a.step_y = 1;
    CHECK(vxCopyImagePatch(in, &r, 0, &a, (void *)y, VX_WRITE_ONLY, VX_MEMORY_TYPE_HOST));
    vx_imagepatch_addressing_t b = {};
    b.dim_x = w; b.dim_y = h; b.stride_x = nv12 ? 2 : 1; b.stride_y = nv12 ? w : w / 2;
    b.scale_x = VX_SCALE_UNITY; b.scale_y = VX_SCALE_UNITY; b.step_x = 2; b.step_y = 2;
    CHECK(vxCopyImagePatch(in, &r, 1, &b, (void *)u, VX_WRITE_ONLY, VX_MEMORY_TYPE_HOST));
    if (!nv12) CHECK(vxCopyImagePatch(in, &r, 2, &b, (void *)v, VX_WRITE_ONLY, VX_MEMORY_TYPE_HOST));
    vx_image out = vxCreateImage(c.ctx, w, h, VX_DF_IMAGE_RGB);
    if (!vxColorConvertNode(c.graph, in, out)) return -1;
    int s = run(c); if (s) return s;
    CHECK(get(out, w, h, rgb, 3));
    return 0;
}

// The planes of an image format as tightly packed host buffers: (width, height, bytes per element,
// addressing step) per plane; chroma planes of 4:2:0 formats are w/2 x h/2 samples.
struct PlaneDesc { int w, h, bpp, step; };

int planes_of(vx_df_image f, int w, int h, PlaneDesc *p) {
    switch (f) {
    case VX_DF_IMAGE_RGB: p[0] = {w, h, 3, 1}; return 1;
    case VX_DF_IMAGE_RGBX: p[0] = {w, h, 4, 1}; return 1;
    case VX_DF_IMAGE_UYVY: case VX_DF_IMAGE_YUYV: p[0] = {w, h, 2, 1}; return 1;
    case VX_DF_IMAGE_IYUV: p[0] = {w, h, 1, 1}; p[1] = {w / 2, h / 2, 1, 2}; p[2] = {w / 2, h / 2, 1, 2}; return 3;
    case VX_DF_IMAGE_NV12: case VX_DF_IMAGE_NV21: p[0] = {w, h, 1, 1}; p[1] = {w / 2, h / 2, 2, 2}; return 2;
    case VX_DF_IMAGE_YUV4: p[0] = {w, h, 1, 1}; p[1] = {w, h, 1, 1}; p[2] = {w, h, 1, 1}; return 3;
    default: return 0;
    }
}

vx_status copy_planes(vx_image img, vx_df_image f, int w, int h, uint8_t *buf, vx_enum usage) {
    PlaneDesc pd[3];
    const int np = planes_of(f, w, h, pd);
    if (!np) return VX_ERROR_INVALID_FORMAT;
    vx_rectangle_t r = {0, 0, (vx_uint32)w, (vx_uint32)h};
    for (int k = 0; k < np; k++) {
        vx_imagepatch_addressing_t a = {};
        a.dim_x = w; a.dim_y = h; a.stride_x = pd[k].bpp; a.stride_y = pd[k].w * pd[k].bpp;
        a.scale_x = VX_SCALE_UNITY; a.scale_y = VX_SCALE_UNITY; a.step_x = pd[k].step; a.step_y = pd[k].step;
        vx_status s = vxCopyImagePatch(img, &r, k, &a, buf, usage, VX_MEMORY_TYPE_HOST);
        if (s != VX_SUCCESS) return s;
        buf += (size_t)pd[k].w * pd[k].h * pd[k].bpp;
    }
    return VX_SUCCESS;
}

// color_convert from any format to any other: `in` and `out` hold every plane of the image, tightly
// packed, one plane after the other (4:2:0 chroma planes at half size, NV12 / NV21 UV interleaved).
int ref_color_convert(uint32_t inFormat, uint32_t outFormat, int w, int h, const uint8_t *in, uint8_t *out) {
    Ctx c;
    vx_image src = vxCreateImage(c.ctx, w, h, (vx_df_image)inFormat);
    vx_image dst = vxCreateImage(c.ctx, w, h, (vx_df_image)outFormat);
    CHECK(copy_planes(src, (vx_df_image)inFormat, w, h, (uint8_t *)in, VX_WRITE_ONLY));
    if (!vxColorConvertNode(c.graph, src, dst)) return -1;
    int s = run(c); if (s) return s;
    CHECK(copy_planes(dst, (vx_df_image)outFormat, w, h, out, VX_READ_ONLY));
    return 0;
}

// color_convert into a virtual image of `yuvFormat` read only by channel_extract(Y): the reference keeps
// only the luma kernel (ColorConvert_Y_RGB / _Y_RGBX), since nothing reads the chroma planes.
int ref_color_luma(uint32_t inFormat, uint32_t yuvFormat, int w, int h, const uint8_t *in, uint8_t *y) {
    Ctx c;
    vx_image src = vxCreateImage(c.ctx, w, h, (vx_df_image)inFormat);
    CHECK(copy_planes(src, (vx_df_image)inFormat, w, h, (uint8_t *)in, VX_WRITE_ONLY));
    vx_image yuv = vxCreateVirtualImage(c.graph, w, h, (vx_df_image)yuvFormat);
    vx_image out = vxCreateImage(c.ctx, w, h, VX_DF_IMAGE_U8);
    if (!vxColorConvertNode(c.graph, src, yuv)) return -1;
    if (!vxChannelExtractNode(c.graph, yuv, VX_CHANNEL_Y, out)) return -1;
    int s = run(c); if (s) return s;
    CHECK(get(out, w, h, y, 1));
    return 0;
}

// RGB -> IYUV (non-virtual output, all three planes): ColorConvert_IYUV_RGB.
int ref_color_iyuv_from_rgb(int w, int h, const uint8_t *rgb, uint8_t *y, uint8_t *u, uint8_t *v) {
    Ctx c;
    vx_image in = make_input(c, w, h, VX_DF_IMAGE_RGB, rgb);
    vx_image out = vxCreateImage(c.ctx, w, h, VX_DF_IMAGE_IYUV);
    if (!vxColorConvertNode(c.graph, in, out)) return -1;
    int s = run(c); if (s) return s;
    vx_rectangle_t r = {0, 0, (vx_uint32)w, (vx_uint32)h};
    vx_imagepatch_addressing_t a = {};
    a.dim_x = w; a.dim_y = h; a.stride_x = 1; a.stride_y = w;
    a.scale_x = VX_SCALE_UNITY; a.scale_y = VX_SCALE_UNITY; a.step_x = 1; a.step_y = 1;
    CHECK(vxCopyImagePatch(out, &r, 0, &a, (void *)y, VX_READ_ONLY, VX_MEMORY_TYPE_HOST));
    vx_imagepatch_addressing_t b = {};
    b.dim_x = w; b.dim_y = h; b.stride_x = 1; b.stride_y = w / 2;
    b.scale_x = VX_SCALE_UNITY; b.scale_y = VX_SCALE_UNITY; b.step_x = 2; b.step_y = 2;
    CHECK(vxCopyImagePatch(out, &r, 1, &b, (void *)u, VX_READ_ONLY, VX_MEMORY_TYPE_HOST));
    CHECK(vxCopyImagePatch(out, &r, 2, &b, (void *)v, VX_READ_ONLY, VX_MEMORY_TYPE_HOST));
    return 0;
}

// halfscale_gaussian with kernel size 1, 3 or 5 (output (w+1)/2 x (h+1)/2).
int ref_halfscale(int w, int h, const uint8_t *src, int ksize, uint8_t *dst) {
    Ctx c;
    vx_image in = make_input(c, w, h, VX_DF_IMAGE_U8, src);
    const int dw = (w + 1) / 2, dh = (h + 1) / 2;
    vx_image out = vxCreateImage(c.ctx, dw, dh, VX_DF_IMAGE_U8);
    if (!vxHalfScaleGaussianNode(c.graph, in, out, ksize)) return -1;
    int s = run(c); if (s) return s;
    CHECK(get(out, dw, dh, dst, 1));
    return 0;
}

static int get_pyramid(vx_pyramid pyr, int levels, uint8_t *out) {
    size_t off = 0;
    for (int l = 0; l < levels; l++) {
        vx_image img = vxGetPyramidLevel(pyr, l);
        vx_uint32 lw = 0, lh = 0;
        vxQueryImage(img, VX_IMAGE_WIDTH, &lw, sizeof(lw));
        vxQueryImage(img, VX_IMAGE_HEIGHT, &lh, sizeof(lh));
        CHECK(get(img, lw, lh, out + off, 1));
        off += (size_t)lw * lh;
        vxReleaseImage(&img);
    }
    return 0;
}

// GaussianPyramid HALF; `out` receives the levels packed back to back.
int ref_pyramid(int w, int h, const uint8_t *src, int levels, uint8_t *out) {
    Ctx c;
    vx_image in = make_input(c, w, h, VX_DF_IMAGE_U8, src);
    vx_pyramid pyr = vxCreatePyramid(c.ctx, levels, VX_SCALE_PYRAMID_HALF, w, h, VX_DF_IMAGE_U8);
    if (!vxGaussianPyramidNode(c.graph, in, pyr)) return -1;
    int s = run(c); if (s) return s;
    return get_pyramid(pyr, levels, out);
}

static vx_threshold make_range(Ctx &c, int lower, int upper) {
    vx_threshold thr = vxCreateThresholdForImage(c.ctx, VX_THRESHOLD_TYPE_RANGE, VX_DF_IMAGE_U8, VX_DF_IMAGE_U8);
    vx_pixel_value_t lo, hi;
    memset(&lo, 0, sizeof(lo)); memset(&hi, 0, sizeof(hi));
    lo.U8 = (vx_uint8)lower; hi.U8 = (vx_uint8)upper;
    vxCopyThresholdRange(thr, &lo, &hi, VX_WRITE_ONLY, VX_MEMORY_TYPE_HOST);
    return thr;
}

// CannyEdgeDetector, gradient size 3, L1 norm.
int ref_canny(int w, int h, const uint8_t *src, int lower, int upper, uint8_t *dst) {
    Ctx c;
    vx_image in = make_input(c, w, h, VX_DF_IMAGE_U8, src);
    vx_image out = vxCreateImage(c.ctx, w, h, VX_DF_IMAGE_U8);
    vx_threshold thr = make_range(c, lower, upper);
    if (!vxCannyEdgeDetectorNode(c.graph, in, thr, 3, VX_NORM_L1, out)) return -1;
    int s = run(c); if (s) return s;
    CHECK(get(out, w, h, dst, 1));
    return 0;
}

// CannyEdgeDetector with any gradient size / norm; the range threshold is created for S16 input so
// that values above 255 fit.  The graph runs `reps` times (the gradient-7 path divides the threshold
// object's values by 4 on every run); *lo_after / *hi_after receive the threshold values afterwards.
int ref_canny_ex(int w, int h, const uint8_t *src, int gsize, int l2, int lower, int upper, int reps, uint8_t *dst,
                 int *lo_after, int *hi_after) {
    Ctx c;
    vx_image in = make_input(c, w, h, VX_DF_IMAGE_U8, src);
    vx_image out = vxCreateImage(c.ctx, w, h, VX_DF_IMAGE_U8);
    vx_threshold thr = vxCreateThresholdForImage(c.ctx, VX_THRESHOLD_TYPE_RANGE, VX_DF_IMAGE_S16, VX_DF_IMAGE_U8);
    vx_pixel_value_t lo, hi;
    memset(&lo, 0, sizeof(lo)); memset(&hi, 0, sizeof(hi));
    lo.S16 = (vx_int16)lower; hi.S16 = (vx_int16)upper;
    CHECK(vxCopyThresholdRange(thr, &lo, &hi, VX_WRITE_ONLY, VX_MEMORY_TYPE_HOST));
    if (!vxCannyEdgeDetectorNode(c.graph, in, thr, gsize, l2 ? VX_NORM_L2 : VX_NORM_L1, out)) return -1;
    CHECK(vxVerifyGraph(c.graph));
    for (int r = 0; r < reps; r++) CHECK(vxProcessGraph(c.graph));
    CHECK(get(out, w, h, dst, 1));
    CHECK(vxCopyThresholdRange(thr, &lo, &hi, VX_READ_ONLY, VX_MEMORY_TYPE_HOST));
    *lo_after = lo.S16;
    *hi_after = hi.S16;
    return 0;
}

static vx_node harris_node(Ctx &c, vx_image in, float k, float thr, float dist, vx_array corners, vx_scalar num) {
    vx_scalar sthr = vxCreateScalar(c.ctx, VX_TYPE_FLOAT32, &thr);
    vx_scalar sdist = vxCreateScalar(c.ctx, VX_TYPE_FLOAT32, &dist);
    vx_scalar sk = vxCreateScalar(c.ctx, VX_TYPE_FLOAT32, &k);
    return vxHarrisCornersNode(c.graph, in, sthr, sdist, sk, 3, 3, corners, num);
}

static int get_corners(vx_array corners, vx_scalar num, vx_keypoint_t *out, int *numitems, long *count) {
    vx_size n = 0, total = 0;
    vxQueryArray(corners, VX_ARRAY_NUMITEMS, &n, sizeof(n));
    if (n) CHECK(vxCopyArrayRange(corners, 0, n, sizeof(vx_keypoint_t), out, VX_READ_ONLY, VX_MEMORY_TYPE_HOST));
    CHECK(vxCopyScalar(num, &total, VX_READ_ONLY, VX_MEMORY_TYPE_HOST));
    *numitems = (int)n;
    *count = (long)total;
    return 0;
}

// HarrisCorners, gradient 3, block 3.  *numitems = items in the array (<= capacity), *count = the
// num_corners scalar (uncapped).
int ref_harris(int w, int h, const uint8_t *src, float k, float thr, float dist, int capacity,
               vx_keypoint_t *out, int *numitems, long *count) {
    Ctx c;
    vx_image in = make_input(c, w, h, VX_DF_IMAGE_U8, src);
    vx_array corners = vxCreateArray(c.ctx, VX_TYPE_KEYPOINT, capacity);
    vx_size zero = 0;
    vx_scalar num = vxCreateScalar(c.ctx, VX_TYPE_SIZE, &zero);
    if (!harris_node(c, in, k, thr, dist, corners, num)) return -1;
    int s = run(c); if (s) return s;
    return get_corners(corners, num, out, numitems, count);
}

// FastCorners with and without non-maximum suppression.
int ref_fast(int w, int h, const uint8_t *src, float thr, int nonmax, int capacity, vx_keypoint_t *out, int *numitems,
             long *count) {
    Ctx c;
    vx_image in = make_input(c, w, h, VX_DF_IMAGE_U8, src);
    vx_array corners = vxCreateArray(c.ctx, VX_TYPE_KEYPOINT, capacity);
    vx_size zero = 0;
    vx_scalar num = vxCreateScalar(c.ctx, VX_TYPE_SIZE, &zero);
    vx_scalar sthr = vxCreateScalar(c.ctx, VX_TYPE_FLOAT32, &thr);
    if (!vxFastCornersNode(c.graph, in, sthr, nonmax ? vx_true_e : vx_false_e, corners, num)) return -1;
    int s = run(c); if (s) return s;
    return get_corners(corners, num, out, numitems, count);
}

// HarrisCorners with gradient size g and block size b.
int ref_harris_ex(int w, int h, const uint8_t *src, float k, float thr, float dist, int g, int b, int capacity,
                  vx_keypoint_t *out, int *numitems, long *count) {
    Ctx c;
    vx_image in = make_input(c, w, h, VX_DF_IMAGE_U8, src);
    vx_array corners = vxCreateArray(c.ctx, VX_TYPE_KEYPOINT, capacity);
    vx_size zero = 0;
    vx_scalar num = vxCreateScalar(c.ctx, VX_TYPE_SIZE, &zero);
    vx_scalar sthr = vxCreateScalar(c.ctx, VX_TYPE_FLOAT32, &thr);
    vx_scalar sdist = vxCreateScalar(c.ctx, VX_TYPE_FLOAT32, &dist);
    vx_scalar sk = vxCreateScalar(c.ctx, VX_TYPE_FLOAT32, &k);
    if (!vxHarrisCornersNode(c.graph, in, sthr, sdist, sk, g, b, corners, num)) return -1;
    int s = run(c); if (s) return s;
    return get_corners(corners, num, out, numitems, count);
}

// The 4K vision graph: RGB -> IYUV (virtual) -> Y (virtual); Y -> ScaleImage (dw x dh, bilinear);
// Y -> GaussianPyramid (levels, HALF); Y -> Canny (3, L1, lower/upper); Y -> Harris (k, thr, dist, 3, 3).
// If time_iters > 0 the graph is processed that many more times; *sec_per_frame gets the mean.
int ref_graph2(int w, int h, const uint8_t *rgb, int dw, int dh, uint8_t *scaled, int levels, uint8_t *pyr_out,
               int lower, int upper, uint8_t *edges, float k, float thr, float dist, int capacity,
               vx_keypoint_t *kp, int *numitems, long *count, int time_iters, double *sec_per_frame) {
    Ctx c;
    vx_image in = make_input(c, w, h, VX_DF_IMAGE_RGB, rgb);
    vx_image iyuv = vxCreateVirtualImage(c.graph, w, h, VX_DF_IMAGE_IYUV);
    vx_image y = vxCreateVirtualImage(c.graph, w, h, VX_DF_IMAGE_U8);
    vx_image sc = vxCreateImage(c.ctx, dw, dh, VX_DF_IMAGE_U8);
    vx_pyramid pyr = vxCreatePyramid(c.ctx, levels, VX_SCALE_PYRAMID_HALF, w, h, VX_DF_IMAGE_U8);
    vx_image ed = vxCreateImage(c.ctx, w, h, VX_DF_IMAGE_U8);
    vx_threshold th = make_range(c, lower, upper);
    vx_array corners = vxCreateArray(c.ctx, VX_TYPE_KEYPOINT, capacity);
    vx_size zero = 0;
    vx_scalar num = vxCreateScalar(c.ctx, VX_TYPE_SIZE, &zero);
    if (!vxColorConvertNode(c.graph, in, iyuv)) return -1;
    if (!vxChannelExtractNode(c.graph, iyuv, VX_CHANNEL_Y, y)) return -1;
    if (!vxScaleImageNode(c.graph, y, sc, VX_INTERPOLATION_BILINEAR)) return -1;
    if (!vxGaussianPyramidNode(c.graph, y, pyr)) return -1;
    if (!vxCannyEdgeDetectorNode(c.graph, y, th, 3, VX_NORM_L1, ed)) return -1;
    if (!harris_node(c, y, k, thr, dist, corners, num)) return -1;
    int s = run(c); if (s) return s;
    CHECK(get(sc, dw, dh, scaled, 1));
    if (get_pyramid(pyr, levels, pyr_out)) return -1;
    CHECK(get(ed, w, h, edges, 1));
    if (get_corners(corners, num, kp, numitems, count)) return -1;
    if (time_iters > 0 && sec_per_frame) {
        auto t0 = std::chrono::steady_clock::now();
        for (int i = 0; i < time_iters; i++) CHECK(vxProcessGraph(c.graph));
        auto t1 = std::chrono::steady_clock::now();
        *sec_per_frame = std::chrono::duration<double>(t1 - t0).count() / time_iters;
    }
    return 0;
}

// WarpAffine (kind 0, m = 2x3 VX matrix, 6 floats) / WarpPerspective (kind 1, 3x3, 9 floats), bilinear;
// border UNDEFINED (constant == 0) or CONSTANT with `border`.
int ref_warp(int kind, int sw, int sh, const uint8_t *src, int dw, int dh, const float *m, int constant, int border,
             uint8_t *dst, int time_iters, double *sec_per_frame) {
    Ctx c;
    vx_image in = make_input(c, sw, sh, VX_DF_IMAGE_U8, src);
    vx_image out = vxCreateImage(c.ctx, dw, dh, VX_DF_IMAGE_U8);
    vx_matrix mat = vxCreateMatrix(c.ctx, VX_TYPE_FLOAT32, kind ? 3 : 2, 3);
    CHECK(vxCopyMatrix(mat, (void *)m, VX_WRITE_ONLY, VX_MEMORY_TYPE_HOST));
    vx_node n = kind ? vxWarpPerspectiveNode(c.graph, in, mat, VX_INTERPOLATION_BILINEAR, out)
                     : vxWarpAffineNode(c.graph, in, mat, VX_INTERPOLATION_BILINEAR, out);
    if (!n) return -1;
    if (constant) {
        vx_border_t b;
        memset(&b, 0, sizeof(b));
        b.mode = VX_BORDER_CONSTANT;
        b.constant_value.U8 = (vx_uint8)border;
        CHECK(vxSetNodeAttribute(n, VX_NODE_BORDER, &b, sizeof(b)));
    }
    int s = run(c); if (s) return s;
    CHECK(get(out, dw, dh, dst, 1));
    if (time_iters > 0 && sec_per_frame) {
        auto t0 = std::chrono::steady_clock::now();
        for (int i = 0; i < time_iters; i++) CHECK(vxProcessGraph(c.graph));
        auto t1 = std::chrono::steady_clock::now();
        *sec_per_frame = std::chrono::duration<double>(t1 - t0).count() / time_iters;
    }
    return 0;
}

// WarpAffine / WarpPerspective with any interpolation; `dst` receives ALIGN16(dw)-byte rows, padding included.
int ref_warp_padded(int kind, int interp, int sw, int sh, const uint8_t *src, int dw, int dh, const float *m,
                    int constant, int border, uint8_t *dst) {
    Ctx c;
    vx_image in = make_input(c, sw, sh, VX_DF_IMAGE_U8, src);
    vx_image out = vxCreateImage(c.ctx, dw, dh, VX_DF_IMAGE_U8);
    vx_matrix mat = vxCreateMatrix(c.ctx, VX_TYPE_FLOAT32, kind ? 3 : 2, 3);
    CHECK(vxCopyMatrix(mat, (void *)m, VX_WRITE_ONLY, VX_MEMORY_TYPE_HOST));
    vx_node n = kind ? vxWarpPerspectiveNode(c.graph, in, mat, (vx_enum)interp, out)
                     : vxWarpAffineNode(c.graph, in, mat, (vx_enum)interp, out);
    if (!n) return -1;
    if (constant) {
        vx_border_t b;
        memset(&b, 0, sizeof(b));
        b.mode = VX_BORDER_CONSTANT;
        b.constant_value.U8 = (vx_uint8)border;
        CHECK(vxSetNodeAttribute(n, VX_NODE_BORDER, &b, sizeof(b)));
    }
    int s = run(c); if (s) return s;
    CHECK(get_padded(out, dw, dh, dst));
    return 0;
}

// Remap with any interpolation from a float map; `dst` receives ALIGN16(dw)-byte rows, padding included.
int ref_remap_padded(int interp, int sw, int sh, const uint8_t *src, int dw, int dh, const float *map, int constant,
                     int border, uint8_t *dst) {
    Ctx c;
    vx_image in = make_input(c, sw, sh, VX_DF_IMAGE_U8, src);
    vx_image out = vxCreateImage(c.ctx, dw, dh, VX_DF_IMAGE_U8);
    vx_remap rm = vxCreateRemap(c.ctx, sw, sh, dw, dh);
    vx_rectangle_t r = {0, 0, (vx_uint32)dw, (vx_uint32)dh};
    CHECK(vxCopyRemapPatch(rm, &r, (vx_size)dw * 8, (void *)map, VX_TYPE_COORDINATES2DF, VX_WRITE_ONLY, VX_MEMORY_TYPE_HOST));
    vx_node n = vxRemapNode(c.graph, in, rm, (vx_enum)interp, out);
    if (!n) return -1;
    if (constant) {
        vx_border_t b;
        memset(&b, 0, sizeof(b));
        b.mode = VX_BORDER_CONSTANT;
        b.constant_value.U8 = (vx_uint8)border;
        CHECK(vxSetNodeAttribute(n, VX_NODE_BORDER, &b, sizeof(b)));
    }
    int s = run(c); if (s) return s;
    CHECK(get_padded(out, dw, dh, dst));
    return 0;
}

// Remap bilinear from a float map (dw*dh {x, y} pairs, set through vxCopyRemapPatch); border as above.
int ref_remap(int sw, int sh, const uint8_t *src, int dw, int dh, const float *map, int constant, int border,
              uint8_t *dst, int time_iters, double *sec_per_frame) {
    Ctx c;
    vx_image in = make_input(c, sw, sh, VX_DF_IMAGE_U8, src);
    vx_image out = vxCreateImage(c.ctx, dw, dh, VX_DF_IMAGE_U8);
    vx_remap rm = vxCreateRemap(c.ctx, sw, sh, dw, dh);
    vx_rectangle_t r = {0, 0, (vx_uint32)dw, (vx_uint32)dh};
    CHECK(vxCopyRemapPatch(rm, &r, (vx_size)dw * 8, (void *)map, VX_TYPE_COORDINATES2DF, VX_WRITE_ONLY, VX_MEMORY_TYPE_HOST));
    vx_node n = vxRemapNode(c.graph, in, rm, VX_INTERPOLATION_BILINEAR, out);
    if (!n) return -1;
    if (constant) {
        vx_border_t b;
        memset(&b, 0, sizeof(b));
        b.mode = VX_BORDER_CONSTANT;
        b.constant_value.U8 = (vx_uint8)border;
        CHECK(vxSetNodeAttribute(n, VX_NODE_BORDER, &b, sizeof(b)));
    }
    int s = run(c); if (s) return s;
    CHECK(get(out, dw, dh, dst, 1));
    if (time_iters > 0 && sec_per_frame) {
        auto t0 = std::chrono::steady_clock::now();
        for (int i = 0; i < time_iters; i++) CHECK(vxProcessGraph(c.graph));
        auto t1 = std::chrono::steady_clock::now();
        *sec_per_frame = std::chrono::duration<double>(t1 - t0).count() / time_iters;
    }
    return 0;
}

// OpticalFlowPyrLK between two images through HALF Gaussian pyramids built in the same graph
// (GaussianPyramid nodes).  old_kp / est_kp: n vx_keypoint_t; new_kp receives n items.  The graph
// is processed `reps` times (the node keeps scratch state between runs); new_kp holds the last run.
int ref_lk(int w, int h, const uint8_t *old_img, const uint8_t *new_img, int levels, int n, const vx_keypoint_t *old_kp,
           const vx_keypoint_t *est_kp, vx_keypoint_t *new_kp, int termination, float epsilon, int num_iterations,
           int use_initial_estimate, int window, int reps, double *sec_per_run) {
    Ctx c;
    vx_image io = make_input(c, w, h, VX_DF_IMAGE_U8, old_img);
    vx_image in = make_input(c, w, h, VX_DF_IMAGE_U8, new_img);
    vx_pyramid po = vxCreatePyramid(c.ctx, levels, VX_SCALE_PYRAMID_HALF, w, h, VX_DF_IMAGE_U8);
    vx_pyramid pn = vxCreatePyramid(c.ctx, levels, VX_SCALE_PYRAMID_HALF, w, h, VX_DF_IMAGE_U8);
    vx_array ao = vxCreateArray(c.ctx, VX_TYPE_KEYPOINT, n);
    vx_array ae = vxCreateArray(c.ctx, VX_TYPE_KEYPOINT, n);
    vx_array an = vxCreateArray(c.ctx, VX_TYPE_KEYPOINT, n);
    CHECK(vxAddArrayItems(ao, n, old_kp, sizeof(vx_keypoint_t)));
    CHECK(vxAddArrayItems(ae, n, est_kp, sizeof(vx_keypoint_t)));
    vx_float32 eps = epsilon;
    vx_uint32 iters = (vx_uint32)num_iterations;
    vx_bool use = use_initial_estimate ? vx_true_e : vx_false_e;
    vx_scalar s_eps = vxCreateScalar(c.ctx, VX_TYPE_FLOAT32, &eps);
    vx_scalar s_it = vxCreateScalar(c.ctx, VX_TYPE_UINT32, &iters);
    vx_scalar s_use = vxCreateScalar(c.ctx, VX_TYPE_BOOL, &use);
    if (!vxGaussianPyramidNode(c.graph, io, po)) return -1;
    if (!vxGaussianPyramidNode(c.graph, in, pn)) return -1;
    if (!vxOpticalFlowPyrLKNode(c.graph, po, pn, ao, ae, an, (vx_enum)termination, s_eps, s_it, s_use, (vx_size)window)) return -1;
    CHECK(vxVerifyGraph(c.graph));
    auto t0 = std::chrono::steady_clock::now();
    for (int r = 0; r < reps; r++) CHECK(vxProcessGraph(c.graph));
    auto t1 = std::chrono::steady_clock::now();
    if (sec_per_run) *sec_per_run = std::chrono::duration<double>(t1 - t0).count() / (reps > 0 ? reps : 1);
    vx_size got = 0;
    CHECK(vxQueryArray(an, VX_ARRAY_NUMITEMS, &got, sizeof(got)));
    if ((int)got != n) return -2;
    CHECK(vxCopyArrayRange(an, 0, n, sizeof(vx_keypoint_t), new_kp, VX_READ_ONLY, VX_MEMORY_TYPE_HOST));
    return 0;
}

}  // extern "C"
