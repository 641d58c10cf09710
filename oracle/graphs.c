/*
 * graphs.c -- whole-graph compositions of the oracle functions (TEST INFRASTRUCTURE).
 * Each follows the kernel sequence the reference's graph optimizer produces for the graph
 * (divide: ago_drama_divide.cpp:509-561, 863-875; merges: ago_drama_remove.cpp:146-182), on
 * planes laid out as the reference allocates them: stride = ALIGN16(width * bpp)
 * (ago_util.cpp:1107) with zeroed padding bytes.  Inputs and outputs of these entry points are
 * packed (stride = width * bpp).
 */
#include "oracle.h"
#include <stdlib.h>
#include <string.h>

#define ALIGN16(x) (((x) + 15) & ~15)

/* a zero-filled plane with the reference's stride */
typedef struct { uint8_t *p; int stride; } plane_t;

static plane_t plane_new(int w, int h, int bpp)
{
    plane_t pl;
    pl.stride = ALIGN16(w * bpp);
    pl.p = (uint8_t *)calloc((size_t)pl.stride * h + 64, 1);
    return pl;
}

static void plane_put(plane_t *pl, int w, int h, int bpp, const void *src)
{
    for (int y = 0; y < h; y++)
        memcpy(pl->p + (size_t)y * pl->stride, (const uint8_t *)src + (size_t)y * w * bpp, (size_t)w * bpp);
}

static void plane_get(const plane_t *pl, int w, int h, int bpp, void *dst)
{
    for (int y = 0; y < h; y++)
        memcpy((uint8_t *)dst + (size_t)y * w * bpp, pl->p + (size_t)y * pl->stride, (size_t)w * bpp);
}

/* One 3x3 U8 filter node on a fresh output (which: 0 gaussian, 1 box, 2 median). */
void oracle_filter3x3(int which, int w, int h, const uint8_t *src, uint8_t *dst)
{
    plane_t in = plane_new(w, h, 1), out = plane_new(w, h, 1);
    plane_put(&in, w, h, 1, src);
    if (which == 0) oracle_gaussian3x3(w, h, out.p, out.stride, in.p, in.stride);
    else if (which == 1) oracle_box3x3(w, h, out.p, out.stride, in.p, in.stride);
    else oracle_median3x3(w, h, out.p, out.stride, in.p, in.stride);
    plane_get(&out, w, h, 1, dst);
    free(in.p); free(out.p);
}

/* Sobel3x3 node on fresh outputs; gx or gy may be NULL. */
void oracle_sobel_node(int w, int h, const uint8_t *src, int16_t *gx, int16_t *gy)
{
    plane_t in = plane_new(w, h, 1), ox = plane_new(w, h, 2), oy = plane_new(w, h, 2);
    plane_put(&in, w, h, 1, src);
    oracle_sobel3x3(w, h, (int16_t *)ox.p, ox.stride, (int16_t *)oy.p, oy.stride, in.p, in.stride);
    if (gx) plane_get(&ox, w, h, 2, gx);
    if (gy) plane_get(&oy, w, h, 2, gy);
    free(in.p); free(ox.p); free(oy.p);
}

/* Gaussian3x3 (virtual) -> Sobel3x3 (gx, gy kept) -> Magnitude + Phase.  gx/gy are graph
 * outputs, so the SOLITARY merge rules (which need virtual gx/gy) do not fire and the
 * reference runs GAUSSIAN_U8_U8_3x3, SOBEL_S16S16_U8_3x3_GXY, MAGNITUDE_S16_S16S16 and
 * PHASE_U8_S16S16. */
void oracle_chain1(int w, int h, const uint8_t *src, int16_t *gx, int16_t *gy, int16_t *mag,
                   uint8_t *phase)
{
    plane_t in = plane_new(w, h, 1), g = plane_new(w, h, 1);
    plane_t ox = plane_new(w, h, 2), oy = plane_new(w, h, 2), om = plane_new(w, h, 2), op = plane_new(w, h, 1);
    plane_put(&in, w, h, 1, src);
    oracle_gaussian3x3(w, h, g.p, g.stride, in.p, in.stride);
    oracle_sobel3x3(w, h, (int16_t *)ox.p, ox.stride, (int16_t *)oy.p, oy.stride, g.p, g.stride);
    oracle_magnitude(w, h, (int16_t *)om.p, om.stride, (int16_t *)ox.p, ox.stride, (int16_t *)oy.p, oy.stride);
    oracle_phase(w, h, op.p, op.stride, (int16_t *)ox.p, ox.stride, (int16_t *)oy.p, oy.stride);
    plane_get(&ox, w, h, 2, gx);
    plane_get(&oy, w, h, 2, gy);
    plane_get(&om, w, h, 2, mag);
    plane_get(&op, w, h, 1, phase);
    free(in.p); free(g.p); free(ox.p); free(oy.p); free(om.p); free(op.p);
}

/* ColorConvert RGB -> IYUV (virtual) -> ChannelExtract Y: the reference keeps only
 * COLOR_CONVERT_Y_RGB (the U/V kernels have no reader and are removed, the channel copy of a
 * virtual plane is removed). */
void oracle_color_y_node(int w, int h, const uint8_t *rgb, uint8_t *y)
{
    plane_t in = plane_new(w, h, 3), out = plane_new(w, h, 1);
    plane_put(&in, w, h, 3, rgb);
    oracle_color_y_rgb(w, h, out.p, out.stride, in.p, in.stride);
    plane_get(&out, w, h, 1, y);
    free(in.p); free(out.p);
}

void oracle_scale_node(int sw, int sh, const uint8_t *src, int dw, int dh, uint8_t *dst)
{
    plane_t in = plane_new(sw, sh, 1), out = plane_new(dw, dh, 1);
    plane_put(&in, sw, sh, 1, src);
    oracle_scale_bilinear(dw, dh, out.p, out.stride, sw, sh, in.p, in.stride);
    plane_get(&out, dw, dh, 1, dst);
    free(in.p); free(out.p);
}

/* Level sizes of a HALF pyramid: ceil(0.5 * previous) (ago_util.cpp:1312-1334). */
void oracle_pyramid_dims(int w, int h, int levels, int *ws, int *hs)
{
    for (int l = 0; l < levels; l++) {
        ws[l] = w; hs[l] = h;
        w = (w + 1) / 2; h = (h + 1) / 2;
    }
}

/* GaussianPyramid HALF: level 0 = channel copy, level l = ScaleGaussianHalf_5x5(level l-1)
 * (ago_drama_divide.cpp:900-922).  `out` receives the packed levels back to back. */
int oracle_pyramid_node(int w, int h, const uint8_t *src, int levels, uint8_t *out)
{
    int ws[32], hs[32];
    oracle_pyramid_dims(w, h, levels, ws, hs);
    plane_t prev = plane_new(w, h, 1);
    plane_put(&prev, w, h, 1, src);
    int rc = 0;
    size_t off = 0;
    plane_get(&prev, w, h, 1, out);
    off += (size_t)w * h;
    for (int l = 1; l < levels; l++) {
        plane_t cur = plane_new(ws[l], hs[l], 1);
        if (oracle_scale_gaussian_half_5x5(ws[l], hs[l], cur.p, cur.stride, ws[l - 1], hs[l - 1], prev.p, prev.stride))
            rc = -1;
        plane_get(&cur, ws[l], hs[l], 1, out + off);
        off += (size_t)ws[l] * hs[l];
        free(prev.p);
        prev = cur;
    }
    free(prev.p);
    return rc;
}

/* CannyEdgeDetector (gradient 3, L1): CANNY_SOBEL_U16_U8_3x3_L1NORM -> CANNY_SUPP_THRESHOLD_U8XY_U16_3x3
 * -> CANNY_EDGE_TRACE_U8_U8XY (ago_drama_divide.cpp:1618-1701, USE_AGO_CANNY_SOBEL_SUPP_THRESHOLD 0). */
void oracle_canny_node(int w, int h, const uint8_t *src, int lower, int upper, uint8_t *dst)
{
    plane_t in = plane_new(w, h, 1), mag = plane_new(w, h, 2), out = plane_new(w, h, 1);
    plane_put(&in, w, h, 1, src);
    oracle_canny_sobel_3x3_l1(w, h, (uint16_t *)mag.p, mag.stride, in.p, in.stride);
    oracle_canny_supp_threshold(w, h, out.p, out.stride, (const uint16_t *)mag.p, mag.stride, lower, upper);
    oracle_canny_trace(w, h, out.p, out.stride);
    plane_get(&out, w, h, 1, dst);
    free(in.p); free(mag.p); free(out.p);
}

/* Canny's U16 gradient plane alone (for kernel-level tests) */
void oracle_canny_sobel_node(int w, int h, const uint8_t *src, uint16_t *dst)
{
    plane_t in = plane_new(w, h, 1), mag = plane_new(w, h, 2);
    plane_put(&in, w, h, 1, src);
    oracle_canny_sobel_3x3_l1(w, h, (uint16_t *)mag.p, mag.stride, in.p, in.stride);
    plane_get(&mag, w, h, 2, dst);
    free(in.p); free(mag.p);
}

/* HarrisCorners (gradient 3, block 3): HARRIS_SOBEL_HG3 -> HARRIS_SCORE_HVC -> NON_MAX_SUPP_XY_ANY_3x3
 * (XYS capacity w*h) -> HARRIS_MERGE_SORT_AND_PICK (ago_drama_divide.cpp:1703-1787).  Returns the
 * uncapped corner count; `hvc` (w*h floats, may be NULL) receives the score plane. */
/* HarrisCorners with any gradient size g and block size b in {3, 5, 7}: HARRIS_SOBEL_HG3_U8_<g>x<g> ->
 * HARRIS_SCORE_HVC_HG3_<b>x<b> -> NON_MAX_SUPP_XY_ANY_3x3 -> HARRIS_MERGE_SORT_AND_PICK_XY_XYS
 * (ago_drama_divide.cpp:1703-1787). */
int oracle_harris_node_ex(int w, int h, const uint8_t *src, int g, int b, float k, float thr, float min_distance,
                          oracle_keypoint_t *out, int capacity, float *hvc)
{
    plane_t in = plane_new(w, h, 1), hg3 = plane_new(w, h, 12), vc = plane_new(w, h, 4);
    plane_put(&in, w, h, 1, src);
    if (g == 3) oracle_harris_sobel_hg3(w, h, (float *)hg3.p, hg3.stride, in.p, in.stride);
    else oracle_harris_sobel_g(w, h, (float *)hg3.p, hg3.stride, in.p, in.stride, g);
    const float norm = oracle_harris_norm_gb(g, b);
    if (b == 3) oracle_harris_score(w, h, (float *)vc.p, vc.stride, (const float *)hg3.p, hg3.stride, k, thr, norm);
    else oracle_harris_score_b(w, h, (float *)vc.p, vc.stride, (const float *)hg3.p, hg3.stride, k, thr, norm, b);
    if (hvc) plane_get(&vc, w, h, 4, hvc);
    int cap = w * h;
    oracle_xys_t *list = (oracle_xys_t *)malloc(sizeof(oracle_xys_t) * (size_t)(cap > 0 ? cap : 1));
    int n = oracle_nonmax_xys(w, h, (const float *)vc.p, vc.stride, list, cap);
    int count = oracle_harris_pick(list, n, min_distance, w, h, out, capacity);
    free(list); free(in.p); free(hg3.p); free(vc.p);
    return count;
}

int oracle_harris_node(int w, int h, const uint8_t *src, float k, float thr, float min_distance,
                       oracle_keypoint_t *out, int capacity, float *hvc)
{
    plane_t in = plane_new(w, h, 1), hg3 = plane_new(w, h, 12), vc = plane_new(w, h, 4);
    plane_put(&in, w, h, 1, src);
    oracle_harris_sobel_hg3(w, h, (float *)hg3.p, hg3.stride, in.p, in.stride);
    oracle_harris_score(w, h, (float *)vc.p, vc.stride, (const float *)hg3.p, hg3.stride, k, thr, oracle_harris_norm(3));
    if (hvc) plane_get(&vc, w, h, 4, hvc);
    int cap = w * h;
    oracle_xys_t *list = (oracle_xys_t *)malloc(sizeof(oracle_xys_t) * (size_t)(cap > 0 ? cap : 1));
    int n = oracle_nonmax_xys(w, h, (const float *)vc.p, vc.stride, list, cap);
    int count = oracle_harris_pick(list, n, min_distance, w, h, out, capacity);
    free(list); free(in.p); free(hg3.p); free(vc.p);
    return count;
}

/* [Gaussian3x3 (virtual) ->] Sobel3x3 (virtual gx, gy) -> Magnitude and/or Phase: the kernel the
 * merge rules choose (ago_drama_remove.cpp:155-182) on fresh outputs.  which: 0 Magnitude + Phase
 * (SOBEL_MAGNITUDE_PHASE_S16U8_U8_3x3), 1 Magnitude only (SOBEL_MAGNITUDE_S16_U8_3x3), 2 Phase only
 * (SOBEL_PHASE_U8_U8_3x3). */
void oracle_sobel_merged_graph(int which, int blur, int w, int h, const uint8_t *src, int16_t *mag, uint8_t *phase)
{
    plane_t in = plane_new(w, h, 1), g = plane_new(w, h, 1), om = plane_new(w, h, 2), op = plane_new(w, h, 1);
    plane_put(&in, w, h, 1, src);
    const plane_t *s = &in;
    if (blur) {
        oracle_gaussian3x3(w, h, g.p, g.stride, in.p, in.stride);
        s = &g;
    }
    if (which == 2) oracle_sobel_phase_merged(w, h, op.p, op.stride, s->p, s->stride);
    else oracle_sobel_magnitude_phase(w, h, (int16_t *)om.p, om.stride, which == 0 ? op.p : NULL, op.stride, s->p, s->stride);
    if (mag && which != 2) plane_get(&om, w, h, 2, mag);
    if (phase && which != 1) plane_get(&op, w, h, 1, phase);
    free(in.p); free(g.p); free(om.p); free(op.p);
}
