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
