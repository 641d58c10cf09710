/*
 * warp.c -- CPU restatement of the reference's bilinear WarpAffine, WarpPerspective and Remap
 * (U8 -> U8, UNDEFINED and CONSTANT borders) and of vxSetRemapPoint's fixed-point map
 * (TEST INFRASTRUCTURE: the checker, never the product).
 *
 * Memory model: see oracle.h.  Source bytes are read LINEARLY at y*stride + x; reads outside
 * [0, h*stride) return 0 (the reference's zero guard).  The reference reads up to one row past the
 * end of the source plane in some cases (the second bilinear row of the last source row); beyond
 * its 64 guard bytes that is heap memory no restatement can know.  Where such a read carries a
 * non-zero weight the pixel is flagged in `unpinned` (parity unpinned, masked in comparisons).
 *
 * The warps write every output row up to ceil(w/4)*4 pixels, as the reference's 4-pixel stores do:
 * pixels w..ceil4(w)-1 land in the row's padding (stride >= ALIGN16(w)).  The reference evaluates
 * those from its per-column coefficient tables (a*x, ... for x < w), whose slots past w are the
 * zeroed gap before the next 16-byte aligned table: padding pixels are computed as column 0 with
 * a table value of +0.0f.  Bilinear remap pads from the following map entries.
 */
#include "oracle.h"
#include <math.h>
#include <string.h>

#define GUARD 64

static inline int rd(const uint8_t *p, long n, long i) { return (i < 0 || i >= n) ? 0 : p[i]; }
static inline int beyond(long n, long i) { return i < -GUARD || i >= n + GUARD; }
static inline uint8_t sat8(int v) { return (uint8_t)(v < 0 ? 0 : v > 255 ? 255 : v); }

/* cvttps_epi32: truncation; out-of-range and NaN give 0x80000000 */
static inline int32_t cvtt(float v)
{
    if (!(v > -2147483904.0f && v < 2147483648.0f)) return INT32_MIN;
    return (int32_t)v;
}

/* cvtps_epi32 with the default rounding (to nearest, ties to even) */
static inline int32_t cvtn(float v)
{
    if (!(v > -2147483904.0f && v < 2147483648.0f)) return INT32_MIN;
    return (int32_t)rintf(v);
}

static inline int32_t imin(int32_t a, int32_t b) { return a < b ? a : b; }
static inline int32_t imax(int32_t a, int32_t b) { return a > b ? a : b; }

/* The corner test of ago_haf_cpu_geometric.cpp:1142-1155: true when some corner of the output maps
 * outside [0,W) x [0,H) (then the masked, clamped loop runs; otherwise the plain loop). */
int oracle_warp_affine_needs_border(int dw, int dh, int sw, int sh, const float m[6])
{
    const float r00 = m[0], r10 = m[1], r01 = m[2], r11 = m[3], c1 = m[4], c2 = m[5];
    const float fw = (float)sw, fh = (float)sh;
    int b = (c1 < 0) | (c2 < 0) | (c1 >= fw) | (c2 >= fh);
    float x1 = r00 * (float)dw + c1, y1 = r10 * (float)dw + c2;
    b |= (x1 < 0) | (y1 < 0) | (x1 >= fw) | (y1 >= fh);
    x1 = r01 * (float)dh + c1;
    y1 = r11 * (float)dh + c2;
    b |= (x1 < 0) | (y1 < 0) | (x1 >= fw) | (y1 >= fh);
    x1 = r00 * (float)dw + r01 * (float)dh + c1;
    y1 = r10 * (float)dw + r11 * (float)dh + c2;
    b |= (x1 < 0) | (y1 < 0) | (x1 >= fw) | (y1 >= fh);
    return b;
}

/* HafCpu_WarpAffine_U8_U8_Bilinear (ago_haf_cpu_geometric.cpp:825-1099) when constant == 0 and
 * HafCpu_WarpAffine_U8_U8_Bilinear_Constant (:1100-1370) when constant != 0.  m is the VX 2x3
 * matrix as stored (ago_affine_matrix_t, matrix[3][2] row-major): xs = m0*x + m2*y + m4,
 * ys = m1*x + m3*y + m5, evaluated as fl(fl(m0*x) + fl(fl(y*m2) + m4)).
 * Border loop (any corner outside): mask = 0 <= xs < W && 0 <= ys < H; xi = trunc(xs),
 * fx = xs - xi; linear offset yi*stride + xi (32-bit wrap) clamped to [0, limit] with limit =
 * H*stride - 1 (CONSTANT) or (H-1)*stride - 1 (UNDEFINED); pixels p0 = [off], p1 = [off+1],
 * p2 = [off+stride], p3 = [off+stride+1];
 *   v = ((p0*(1-fx))*(1-fy) + (p1*fx)*(1-fy)) + ((p2*(1-fx))*fy + (p3*fx)*fy), rounded to nearest
 * even and saturated; outside the mask the border value (UNDEFINED: 0).  Plain loop: the same
 * without mask and clamp.
 * Pinned by: bit-exact against oracle/_ref (both borders, both loops, pixels flagged unpinned
 * masked); identity and integer-translation closed forms. */
void oracle_warp_affine_bilinear(int dw, int dh, uint8_t *dst, int dst_stride, int sw, int sh, const uint8_t *src,
                                 int src_stride, const float m[6], int constant, int border, uint8_t *unpinned,
                                 int unpinned_stride)
{
    const float r00 = m[0], r10 = m[1], r01 = m[2], r11 = m[3], c1 = m[4], c2 = m[5];
    const long n = (long)sh * src_stride;
    const int bb = oracle_warp_affine_needs_border(dw, dh, sw, sh, m);
    const int32_t limit = constant ? (int32_t)((uint32_t)sh * (uint32_t)src_stride - 1u)
                                   : (int32_t)((uint32_t)(sh - 1) * (uint32_t)src_stride - 1u);
    const float fw = (float)sw, fh = (float)sh;
    const int wq = (dw + 3) & ~3;
    for (int y = 0; y < dh; y++) {
        const float xd = (float)y * r01 + c1, yd = (float)y * r11 + c2;
        for (int x = 0; x < wq; x++) {
            const float tx = x < dw ? r00 * (float)x : 0.0f, ty = x < dw ? r10 * (float)x : 0.0f;
            const float xs = tx + xd, ys = ty + yd;
            const int inside = xs >= 0.0f && xs < fw && ys >= 0.0f && ys < fh;
            const int32_t xi = cvtt(xs), yi = cvtt(ys);
            const float fx = xs - (float)xi, fy = ys - (float)yi;
            int32_t off = (int32_t)((uint32_t)yi * (uint32_t)src_stride + (uint32_t)xi);
            if (bb) off = imax(imin(off, limit), 0);
            const long o = off;
            const float p0 = (float)rd(src, n, o), p1 = (float)rd(src, n, o + 1);
            const float p2 = (float)rd(src, n, o + src_stride), p3 = (float)rd(src, n, o + src_stride + 1);
            const float omx = 1.0f - fx, omy = 1.0f - fy;
            const float a = (p0 * omx) * omy, b = (p1 * fx) * omy, c = (p2 * omx) * fy, d = (p3 * fx) * fy;
            const float v = (a + b) + (c + d);
            const int keep = !bb || inside;
            dst[(long)y * dst_stride + x] = keep ? sat8(cvtn(v)) : (uint8_t)(constant ? border : 0);
            if (unpinned && x < dw)
                unpinned[(long)y * unpinned_stride + x] =
                    keep && ((beyond(n, o + 1) && fx != 0.0f) || (beyond(n, o + src_stride) && fy != 0.0f) ||
                             (beyond(n, o + src_stride + 1) && fx != 0.0f && fy != 0.0f) || beyond(n, o));
        }
    }
}

/* HafCpu_WarpPerspective_U8_U8_Bilinear_Constant (ago_haf_cpu_geometric.cpp:1579-1873; the
 * UNDEFINED variant :1560-1577 calls it with border 0).  m is the VX 3x3 matrix as stored
 * (matrix[3][3] row-major; a=m0 d=m1 g=m2 b=m3 e=m4 h=m5 c=m6 f=m7 i=m8):
 *   z = 1 / fl(g*x + fl(y*h + i)); xs = fl(a*x + fl(y*b + c)) * z; ys = fl(d*x + fl(y*e + f)) * z;
 *   xi = trunc(xs) - (xs < 0) (so -1.0 gives -2, as the reference does), fx = xs - xi;
 *   mask  = 0 <= xi < W && 0 <= yi < H  selects p0, p2 (else border);
 *   mask1 = -1 <= xi < W && -1 <= yi < H selects p1, p3 (else border);
 *   offset yi*stride + xi clamped to [0, H*stride - 1];
 *   v = (((p0*(1-fx))*(1-fy) + (p1*fx)*(1-fy)) + (p2*(1-fx))*fy) + (p3*fx)*fy, TRUNCATED, saturated.
 * The corner test is disabled in the reference (:1628-1640, bBoder = 1): the masked loop always runs.
 * Pinned by: bit-exact against oracle/_ref (pixels flagged unpinned masked); the identity matrix
 * reproduces the source (x + 0 weights). */
void oracle_warp_perspective_bilinear(int dw, int dh, uint8_t *dst, int dst_stride, int sw, int sh, const uint8_t *src,
                                      int src_stride, const float m[9], int border, uint8_t *unpinned, int unpinned_stride)
{
    const float a = m[0], d = m[1], g = m[2], b = m[3], e = m[4], hh = m[5], c = m[6], f = m[7], ii = m[8];
    const long n = (long)sh * src_stride;
    const int32_t limit = (int32_t)((uint32_t)sh * (uint32_t)src_stride - 1u);
    const int wq = (dw + 3) & ~3;
    for (int y = 0; y < dh; y++) {
        const float xd = (float)y * b + c, yd = (float)y * e + f, zd = (float)y * hh + ii;
        for (int x = 0; x < wq; x++) {
            const float ax = x < dw ? a * (float)x : 0.0f, dx = x < dw ? d * (float)x : 0.0f, gx = x < dw ? g * (float)x : 0.0f;
            const float z = 1.0f / (gx + zd);
            const float xs = (ax + xd) * z, ys = (dx + yd) * z;
            const int32_t xi = (int32_t)((uint32_t)cvtt(xs) - (uint32_t)(xs < 0.0f));
            const int32_t yi = (int32_t)((uint32_t)cvtt(ys) - (uint32_t)(ys < 0.0f));
            const int mask = xi >= 0 && xi < sw && yi >= 0 && yi < sh;
            const int mask1 = xi >= -1 && xi < sw && yi >= -1 && yi < sh;
            const float fx = xs - (float)xi, fy = ys - (float)yi;
            int32_t off = (int32_t)((uint32_t)yi * (uint32_t)src_stride + (uint32_t)xi);
            off = imax(imin(off, limit), 0);
            const long o = off;
            const float p0 = (float)(mask ? rd(src, n, o) : border), p1 = (float)(mask1 ? rd(src, n, o + 1) : border);
            const float p2 = (float)(mask ? rd(src, n, o + src_stride) : border);
            const float p3 = (float)(mask1 ? rd(src, n, o + src_stride + 1) : border);
            const float omx = 1.0f - fx, omy = 1.0f - fy;
            float v = (p0 * omx) * omy;
            v = v + (p1 * fx) * omy;
            v = v + (p2 * omx) * fy;
            v = v + (p3 * fx) * fy;
            dst[(long)y * dst_stride + x] = sat8(cvtt(v));
            if (unpinned && x < dw)
                unpinned[(long)y * unpinned_stride + x] =
                    (mask1 && beyond(n, o + 1) && fx != 0.0f) || (mask && beyond(n, o + src_stride) && fy != 0.0f) ||
                    (mask1 && beyond(n, o + src_stride + 1) && fx != 0.0f && fy != 0.0f);
        }
    }
}

/* vxSetRemapPoint (vx_api.cpp:7865-7894) for source images below 8192 pixels (3 fractional bits,
 * ago_util.cpp:2450-2459): x = (u16)(sx*8 + 0.5f), y likewise, and (0xffff, 0xffff) when
 * sx < 0, sy < 0, sx >= W-1 or sy >= H-1.  map holds dw*dh {x, y} float pairs; fixed receives
 * dw*dh {x, y} u16 pairs.  Pinned by: the GPU conversion and the remap outputs against oracle/_ref. */
void oracle_remap_fixed(int dw, int dh, int sw, int sh, const float *map, uint16_t *fixed)
{
    for (long i = 0; i < (long)dw * dh; i++) {
        const float sx = map[2 * i], sy = map[2 * i + 1];
        uint16_t fx = (uint16_t)(int)(sx * 8.0f + 0.5f), fy = (uint16_t)(int)(sy * 8.0f + 0.5f);
        if (sx < 0.0f || sy < 0.0f || sx >= (float)(sw - 1) || sy >= (float)(sh - 1)) fx = fy = 0xffff;
        fixed[2 * i] = fx;
        fixed[2 * i + 1] = fy;
    }
}

/* HafCpu_Remap_U8_U8_Bilinear (ago_haf_cpu_geometric.cpp:249-370, constant == 0) and
 * HafCpu_Remap_U8_U8_Bilinear_Constant (:372-504, constant != 0).  Each 16-bit map lane equal to
 * 0xffff is replaced by 8 (UNDEFINED: the point (1,1) with zero fraction) or 0 (CONSTANT).
 * fx = v & 7, xi = v >> 3 (rows alike); offset yi*stride + xi (no clamp);
 *   out = (p00*(8-fx)*(8-fy) + p10*fx*(8-fy) + p01*(8-fx)*fy + p11*fx*fy + 32) >> 6.
 * CONSTANT: the result is replaced by the border value when the X lane was 0xffff (the reference
 * masks the 32-bit result with the 16-bit compare mask, so only the x lane counts).  The map has
 * dw entries per row (stride dw*4 bytes); pixels w..ceil4(w)-1 of a row use the next row's
 * entries; past the last entry the map reads as (0, 0), the zeroed tail of its allocation.
 * Pinned by: bit-exact against oracle/_ref; closed form for integer maps (fraction 0 copies the
 * source pixel). */
void oracle_remap_bilinear(int dw, int dh, uint8_t *dst, int dst_stride, int sw, int sh, const uint8_t *src,
                           int src_stride, const uint16_t *fixed, int constant, int border)
{
    const long n = (long)sh * src_stride;
    const long entries = (long)dw * dh;
    const int wq = (dw + 3) & ~3;
    (void)sw;
    for (int y = 0; y < dh; y++) {
        for (int x = 0; x < wq; x++) {
            const long k = (long)y * dw + x;
            uint32_t mx = k < entries ? fixed[2 * k] : 0, my = k < entries ? fixed[2 * k + 1] : 0;
            const int xbad = mx == 0xffff;
            if (mx == 0xffff) mx = constant ? 0 : 8;
            if (my == 0xffff) my = constant ? 0 : 8;
            const int fx = (int)(mx & 7), fy = (int)(my & 7);
            const long o = (long)(my >> 3) * src_stride + (long)(mx >> 3);
            const int p00 = rd(src, n, o), p10 = rd(src, n, o + 1), p01 = rd(src, n, o + src_stride),
                      p11 = rd(src, n, o + src_stride + 1);
            int v = (p00 * (8 - fx) * (8 - fy) + p10 * fx * (8 - fy) + p01 * (8 - fx) * fy + p11 * fx * fy + 32) >> 6;
            if (constant && xbad) v = border;
            dst[(long)y * dst_stride + x] = sat8(v);
        }
    }
}

/* HafCpu_WarpAffine_U8_U8_Nearest (ago_haf_cpu_geometric.cpp:508-656; border 0) and
 * HafCpu_WarpAffine_U8_U8_Nearest_Constant (:658-823).  xs, ys and the corner test as for bilinear;
 * xi = trunc(xs), yi = trunc(ys), offset yi*stride + xi (32-bit wrap).
 * Masked loop (some corner outside): mask = 0 <= xs < W && 0 <= ys < H (and, for CONSTANT, inside
 * the input's valid rectangle valid = {start_x, start_y, end_x, end_y}); offset clamped to
 * [0, H*stride - 1]; the border value outside the mask.  Plain loop: the source byte at the offset.
 * The reference's y-coefficient table starts right after the dw x-coefficients, so its 16-byte
 * loads fault unless dw % 4 == 0; for other widths this restatement evaluates every column by the
 * same formula and the padding pixels as column 0 ("parity unpinned" for dw % 4 != 0).
 * Pinned by: bit-exact against oracle/_ref (widths multiple of 4, both loops, both borders);
 * identity and integer translations copy the source. */
void oracle_warp_affine_nearest(int dw, int dh, uint8_t *dst, int dst_stride, int sw, int sh, const uint8_t *src,
                                int src_stride, const float m[6], int border, const int valid[4])
{
    const float r00 = m[0], r10 = m[1], r01 = m[2], r11 = m[3], c1 = m[4], c2 = m[5];
    const long n = (long)sh * src_stride;
    const int bb = oracle_warp_affine_needs_border(dw, dh, sw, sh, m);
    const int32_t limit = (int32_t)((uint32_t)sh * (uint32_t)src_stride - 1u);
    const float fw = (float)sw, fh = (float)sh;
    const float vl = (float)valid[0], vt = (float)valid[1], vr = (float)valid[2], vb = (float)valid[3];
    const int wq = (dw + 3) & ~3;
    for (int y = 0; y < dh; y++) {
        const float xd = (float)y * r01 + c1, yd = (float)y * r11 + c2;
        for (int x = 0; x < wq; x++) {
            const float tx = x < dw ? r00 * (float)x : 0.0f, ty = x < dw ? r10 * (float)x : 0.0f;
            const float xs = tx + xd, ys = ty + yd;
            const int32_t xi = cvtt(xs), yi = cvtt(ys);
            int32_t off = (int32_t)((uint32_t)yi * (uint32_t)src_stride + (uint32_t)xi);
            int v;
            if (bb) {
                const int mask = xs >= 0.0f && xs < fw && ys >= 0.0f && ys < fh && xs >= vl && xs < vr && ys >= vt && ys < vb;
                off = imax(imin(off, limit), 0);
                v = mask ? rd(src, n, off) : border;
            } else {
                v = rd(src, n, off);
            }
            dst[(long)y * dst_stride + x] = (uint8_t)v;
        }
    }
}

/* HafCpu_WarpPerspective_U8_U8_Nearest_Constant (ago_haf_cpu_geometric.cpp:1406-1558; the UNDEFINED
 * form :1372-1385 calls it with border 0).  z = 1 / fl(g*x + fl(y*h + i)), xs = fl(a*x + fl(y*b + c))*z,
 * ys = fl(d*x + fl(y*e + f))*z (matrix as for bilinear); always the masked loop:
 * mask = 0 <= xs < W && 0 <= ys < H; offset trunc(ys)*stride + trunc(xs) clamped to [0, H*stride - 1];
 * the border value outside the mask.  Padding pixels: column 0 (zeroed table slots).
 * Pinned by: bit-exact against oracle/_ref, padding included; the identity matrix copies the source. */
void oracle_warp_perspective_nearest(int dw, int dh, uint8_t *dst, int dst_stride, int sw, int sh, const uint8_t *src,
                                     int src_stride, const float m[9], int border)
{
    const float a = m[0], d = m[1], g = m[2], b = m[3], e = m[4], hh = m[5], c = m[6], f = m[7], ii = m[8];
    const long n = (long)sh * src_stride;
    const int32_t limit = (int32_t)((uint32_t)sh * (uint32_t)src_stride - 1u);
    const float fw = (float)sw, fh = (float)sh;
    const int wq = (dw + 3) & ~3;
    for (int y = 0; y < dh; y++) {
        const float xd = (float)y * b + c, yd = (float)y * e + f, zd = (float)y * hh + ii;
        for (int x = 0; x < wq; x++) {
            const float ax = x < dw ? a * (float)x : 0.0f, dx = x < dw ? d * (float)x : 0.0f, gx = x < dw ? g * (float)x : 0.0f;
            const float z = 1.0f / (gx + zd);
            const float xs = (ax + xd) * z, ys = (dx + yd) * z;
            const int mask = xs >= 0.0f && xs < fw && ys >= 0.0f && ys < fh;
            const int32_t xi = cvtt(xs), yi = cvtt(ys);
            int32_t off = (int32_t)((uint32_t)yi * (uint32_t)src_stride + (uint32_t)xi);
            off = imax(imin(off, limit), 0);
            dst[(long)y * dst_stride + x] = (uint8_t)(mask ? rd(src, n, off) : border);
        }
    }
}

/* HafCpu_Remap_U8_U8_Nearest (ago_haf_cpu_geometric.cpp:72-154, constant == 0) and
 * HafCpu_Remap_U8_U8_Nearest_Constant (:156-247, constant != 0) over vxSetRemapPoint's 13.3 map.
 * Columns below dw & ~3 (the 4-pixel body), per 16-bit lane v: coordinate (v >> 3) + ((v & 7) > 3);
 *   UNDEFINED: a lane of 0xffff reads as 7, i.e. coordinate 1; the source byte at y*stride + x;
 *   CONSTANT: a lane of 0xffff reads as 0, and the pixel is the border value when either lane was
 *   0xffff or the offset exceeds H*stride - 1.
 * The last dw % 4 columns (byte loop), per lane taken as a signed 16-bit s: coordinate
 *   (s >> 3) + ((s & 7) >> 2) unless s == -1, which gives 0 (UNDEFINED) or the border value itself
 *   as a coordinate (CONSTANT); the source byte at y*stride + x, no bounds test (offsets outside the
 *   plane read 0 here and are flagged in `unpinned`).
 * No padding bytes are written.
 * Pinned by: bit-exact against oracle/_ref (both borders, body and tail columns); integer maps copy
 * the source pixel. */
void oracle_remap_nearest(int dw, int dh, uint8_t *dst, int dst_stride, int sw, int sh, const uint8_t *src,
                          int src_stride, const uint16_t *fixed, int constant, int border, uint8_t *unpinned,
                          int unpinned_stride)
{
    const long n = (long)sh * src_stride;
    const int32_t limit = (int32_t)((uint32_t)sh * (uint32_t)src_stride - 1u);
    const int body = dw & ~3;
    (void)sw;
    for (int y = 0; y < dh; y++) {
        for (int x = 0; x < dw; x++) {
            const long k = (long)y * dw + x;
            uint32_t mx = fixed[2 * k], my = fixed[2 * k + 1];
            int v, far = 0;
            if (x < body) {
                const int bad = mx == 0xffff || my == 0xffff;
                if (mx == 0xffff) mx = constant ? 0 : 7;
                if (my == 0xffff) my = constant ? 0 : 7;
                const uint32_t xi = (mx >> 3) + ((mx & 7) > 3), yi = (my >> 3) + ((my & 7) > 3);
                const int32_t off = (int32_t)(yi * (uint32_t)src_stride + xi);
                if (constant && (bad || off > limit)) {
                    v = border;
                } else {
                    v = rd(src, n, off);
                    far = beyond(n, off);
                }
            } else {
                const int16_t sx = (int16_t)mx, sy = (int16_t)my;
                const int xi = sx != -1 ? (sx >> 3) + ((sx & 7) >> 2) : (constant ? border : 0);
                const int yi = sy != -1 ? (sy >> 3) + ((sy & 7) >> 2) : (constant ? border : 0);
                const uint32_t off = (uint32_t)yi * (uint32_t)src_stride + (uint32_t)xi;
                v = rd(src, n, (long)off);
                far = beyond(n, (long)off);
            }
            dst[(long)y * dst_stride + x] = (uint8_t)v;
            if (unpinned) unpinned[(long)y * unpinned_stride + x] = (uint8_t)far;
        }
    }
}

/* ---- stitch: Loom's merge with its camera-selection maps ----------------------------------------
 * amd_loomsl/kernels/merge.cpp:176-385 generates an OpenCL kernel; Loom has no CPU form, so this is a
 * restatement of that generated source, "parity unpinned" against any reference output and pinned by
 * closed forms in the tests (a selected single camera is copied, 31 leaves the output untouched,
 * complementary weights over equal images copy them, the blend-count thresholds of camId0).
 * Per group of 8 pixels: camId0 < 31 copies that camera; 31 writes nothing; > 31 blends the cameras
 * listed in camId1 (5-bit ids at bits 0, 5, 10) and camId2 (same): two always, a third above 128, camId2's
 * ids above 129, 130, 131 (merge.cpp:248-290).  w = (float)weight * 0.003922f (1/255 printed by "%f" into
 * the kernel source, :226); per channel fa = ((0 + w0*p0) + w1*p1) + ..., each step rounded (mad taken
 * unfused); out = round-to-nearest-even, saturated (amd_pack).  Ids >= count count as 31. */
static const float LOOM_FACTOR = 0.003922f;

static uint8_t loom_pack(float v)
{
    float r = nearbyintf(v);
    if (r < 0.0f) r = 0.0f;
    if (r > 255.0f) r = 255.0f;
    return (uint8_t)r;
}

/* fetch(ctx, camera, x, y, rgb) */
typedef void (*loom_fetch_t)(const void *ctx, int cam, int x, int y, int rgb[3]);

static int loom_px(int x, int y, int count, const uint8_t *wt, int wt_stride, int h, const uint8_t *id0, int id0_stride,
                   const uint16_t *id1, int id1_stride, const uint16_t *id2, int id2_stride, loom_fetch_t fetch, const void *ctx,
                   uint8_t out[3])
{
    const int g = x >> 3;
    const unsigned sel = id0[(long)y * id0_stride + g];
    int p[3];
    if (sel == 31) return 0;
    if (sel < 31) {
        if ((int)sel >= count) return 0;
        fetch(ctx, (int)sel, x, y, p);
        for (int c = 0; c < 3; c++) out[c] = (uint8_t)p[c];
        return 1;
    }
    const unsigned s1 = *(const uint16_t *)((const uint8_t *)id1 + (long)y * id1_stride + 2L * g);
    const unsigned s2 = sel > 129 ? *(const uint16_t *)((const uint8_t *)id2 + (long)y * id2_stride + 2L * g) : 0;
    const int n = sel > 131 ? 6 : sel > 130 ? 5 : sel > 129 ? 4 : sel > 128 ? 3 : 2;
    float fa[3] = {0.0f, 0.0f, 0.0f};
    for (int k = 0; k < n; k++) {
        const unsigned id = ((k < 3 ? s1 : s2) >> (5 * (k % 3))) & 31u;
        float w = 0.0f;
        p[0] = p[1] = p[2] = 0;
        if (id < 31 && (int)id < count) {
            fetch(ctx, (int)id, x, y, p);
            w = (float)wt[((long)id * h + y) * wt_stride + x] * LOOM_FACTOR;
        }
        for (int c = 0; c < 3; c++) {
            const float prod = w * (float)p[c];
            fa[c] = fa[c] + prod;
        }
    }
    for (int c = 0; c < 3; c++) out[c] = loom_pack(fa[c]);
    return 1;
}

typedef struct { const uint8_t *img; int stride, h; } stacked_ctx_t;

static void fetch_stacked(const void *ctx, int cam, int x, int y, int rgb[3])
{
    const stacked_ctx_t *s = (const stacked_ctx_t *)ctx;
    const uint8_t *q = s->img + ((long)cam * s->h + y) * s->stride + 3L * x;
    rgb[0] = q[0]; rgb[1] = q[1]; rgb[2] = q[2];
}

void oracle_stitch_loom(int w, int h, uint8_t *dst, int dst_stride, int count, const uint8_t *stacked, int stacked_stride,
                        const uint8_t *wt, int wt_stride, const uint8_t *id0, int id0_stride, const uint16_t *id1, int id1_stride,
                        const uint16_t *id2, int id2_stride)
{
    stacked_ctx_t s = {stacked, stacked_stride, h};
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            uint8_t o[3];
            if (loom_px(x, y, count, wt, wt_stride, h, id0, id0_stride, id1, id1_stride, id2, id2_stride, fetch_stacked, &s, o))
                for (int c = 0; c < 3; c++) dst[(long)y * dst_stride + 3L * x + c] = o[c];
        }
}

/* Camera `cam` sampled at output (x, y) through its 13.3 table (rows cam*h ..) as
 * oracle_remap_bilinear with a constant border of 0 does for each channel of the packed RGB image
 * (rows y0 .. y1-1 of the output; the others are left as they are). */
typedef struct { const uint8_t *const *cams; int cam_stride, sw, sh, w, h; const uint16_t *fixed; } remap_ctx_t;

static void fetch_remap(const void *ctx, int cam, int x, int y, int rgb[3])
{
    const remap_ctx_t *r = (const remap_ctx_t *)ctx;
    const long k = ((long)cam * r->h + y) * r->w + x;
    unsigned mx = r->fixed[2 * k], my = r->fixed[2 * k + 1];
    const int xbad = mx == 0xffff;
    if (mx == 0xffff) mx = 0;
    if (my == 0xffff) my = 0;
    const int fx = (int)(mx & 7), fy = (int)(my & 7);
    const uint8_t *src = r->cams[cam];
    const long n = (long)r->sh * r->cam_stride;
    for (int c = 0; c < 3; c++) {
        const long o = (long)(my >> 3) * r->cam_stride + 3L * (long)(mx >> 3) + c;
        const int p00 = rd(src, n, o), p10 = rd(src, n, o + 3), p01 = rd(src, n, o + r->cam_stride),
                  p11 = rd(src, n, o + r->cam_stride + 3);
        int v = (p00 * (8 - fx) * (8 - fy) + p10 * fx * (8 - fy) + p01 * (8 - fx) * fy + p11 * fx * fy + 32) >> 6;
        if (xbad) v = 0;
        rgb[c] = sat8(v);
    }
}

void oracle_stitch_loom_remap(int w, int h, int y0, int y1, uint8_t *dst, int dst_stride, int count, const uint8_t *const *cams,
                              int cam_stride, int sw, int sh, const uint16_t *fixed, const uint8_t *wt, int wt_stride, const uint8_t *id0,
                              int id0_stride, const uint16_t *id1, int id1_stride, const uint16_t *id2, int id2_stride)
{
    remap_ctx_t r = {cams, cam_stride, sw, sh, w, h, fixed};
    for (int y = y0; y < y1; y++)
        for (int x = 0; x < w; x++) {
            uint8_t o[3];
            if (loom_px(x, y, count, wt, wt_stride, h, id0, id0_stride, id1, id1_stride, id2, id2_stride, fetch_remap, &r, o))
                for (int c = 0; c < 3; c++) dst[(long)y * dst_stride + 3L * x + c] = o[c];
        }
}
