// warp.hip -- bilinear and nearest-neighbour WarpAffine, WarpPerspective and Remap (U8 -> U8,
// UNDEFINED and CONSTANT borders) and the float -> fixed-point remap table conversion, for gfx950.
//
// Each thread produces four consecutive output pixels and writes them with one 32-bit store, the
// reference's store unit: warp rows are written up to ceil(w/4)*4 pixels, so pixels w..ceil4(w)-1
// land in the row padding with the values the reference's SSE loops leave there (a 3x3 node reading
// the output with linear-address neighbours then sees the same bytes).  The reference evaluates the
// warps from per-column coefficient tables (a*x for x < w) whose slots past w are zero, so a padding
// pixel is column 0's with a coefficient of +0.0f; bilinear remap pads from the next map entries;
// nearest remap writes its last w % 4 pixels one byte at a time and no padding.  The source is read at
// linear byte addresses; addresses outside [0, h*stride) read 0, the reference's zero guard.
// All float arithmetic is single precision, uncontracted and in the reference's order; the
// x86 conversion semantics (truncation / round-to-nearest-even, 0x80000000 for values out of int32
// range) are reproduced explicitly because the GPU's conversions saturate instead.
#include "mvx_common.hpp"

namespace mvx {

namespace {

MVX_DEV uint32_t ldg8(const uint8_t *p, uint64_t n, int64_t i) { return (i < 0 || (uint64_t)i >= n) ? 0u : (uint32_t)p[i]; }

// _mm_cvttps_epi32
MVX_DEV int32_t cvtt(float v)
{
    if (!(v > -2147483904.0f && v < 2147483648.0f)) return INT32_MIN;
    return (int32_t)v;
}

// _mm_cvtps_epi32 (MXCSR default: nearest, ties to even)
MVX_DEV int32_t cvtn(float v)
{
    if (!(v > -2147483904.0f && v < 2147483648.0f)) return INT32_MIN;
    return __float2int_rn(v);
}

MVX_DEV uint32_t sat8(int32_t v) { return (uint32_t)min(max(v, 0), 255); }

struct Affine { float r00, r10, r01, r11, c1, c2; };
struct Persp { float a, d, g, b, e, h, c, f, i; };

}  // namespace

// HafCpu_WarpAffine_U8_U8_Bilinear[_Constant] (ago_haf_cpu_geometric.cpp:825-1370).
// BORDER: the corner test failed, run the masked loop with the offset clamped to [0, limit]
// (limit = H*stride-1 for CONSTANT, (H-1)*stride-1 for UNDEFINED); otherwise the plain loop.
template <bool BORDER>
__global__ void __launch_bounds__(256)
k_warp_affine(uint32_t dw, uint32_t dh, uint8_t *__restrict__ dst, uint32_t dstStride, uint32_t sw, uint32_t sh,
              const uint8_t *__restrict__ src, uint32_t srcStride, Affine m, int32_t limit, uint32_t border)
{
    const uint32_t x0 = (blockIdx.x * blockDim.x + threadIdx.x) * 4;
    const uint32_t y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x0 >= dw || y >= dh) return;
    const uint64_t n = (uint64_t)sh * srcStride;
    const float fw = (float)sw, fh = (float)sh;
    const float xd = (float)y * m.r01 + m.c1, yd = (float)y * m.r11 + m.c2;
    uint32_t word = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const bool col = x0 + k < dw;
        const float xs = (col ? m.r00 * (float)(x0 + k) : 0.0f) + xd, ys = (col ? m.r10 * (float)(x0 + k) : 0.0f) + yd;
        const int32_t xi = cvtt(xs), yi = cvtt(ys);
        const float fx = xs - (float)xi, fy = ys - (float)yi;
        int32_t off = (int32_t)((uint32_t)yi * srcStride + (uint32_t)xi);
        if (BORDER) off = max(min(off, limit), 0);
        const float p0 = (float)ldg8(src, n, off), p1 = (float)ldg8(src, n, (int64_t)off + 1);
        const float p2 = (float)ldg8(src, n, (int64_t)off + srcStride), p3 = (float)ldg8(src, n, (int64_t)off + srcStride + 1);
        const float omx = 1.0f - fx, omy = 1.0f - fy;
        const float a = (p0 * omx) * omy, b = (p1 * fx) * omy, c = (p2 * omx) * fy, d = (p3 * fx) * fy;
        const float v = (a + b) + (c + d);
        uint32_t px = sat8(cvtn(v));
        if (BORDER && !(xs >= 0.0f && xs < fw && ys >= 0.0f && ys < fh)) px = border;
        word |= px << (8 * k);
    }
    *reinterpret_cast<uint32_t *>(dst + (uint64_t)y * dstStride + x0) = word;
}

// HafCpu_WarpPerspective_U8_U8_Bilinear_Constant (ago_haf_cpu_geometric.cpp:1579-1873); the
// UNDEFINED variant is the same with border 0 (:1560-1577).
__global__ void __launch_bounds__(256)
k_warp_perspective(uint32_t dw, uint32_t dh, uint8_t *__restrict__ dst, uint32_t dstStride, uint32_t sw, uint32_t sh,
                   const uint8_t *__restrict__ src, uint32_t srcStride, Persp m, uint32_t border)
{
    const uint32_t x0 = (blockIdx.x * blockDim.x + threadIdx.x) * 4;
    const uint32_t y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x0 >= dw || y >= dh) return;
    const uint64_t n = (uint64_t)sh * srcStride;
    const int32_t limit = (int32_t)(sh * srcStride - 1u);
    const float xd = (float)y * m.b + m.c, yd = (float)y * m.e + m.f, zd = (float)y * m.h + m.i;
    const float bf = (float)border;
    uint32_t word = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const bool col = x0 + k < dw;
        const float xf = (float)(x0 + k);
        const float z = 1.0f / ((col ? m.g * xf : 0.0f) + zd);
        const float xs = ((col ? m.a * xf : 0.0f) + xd) * z, ys = ((col ? m.d * xf : 0.0f) + yd) * z;
        const int32_t xi = (int32_t)((uint32_t)cvtt(xs) - (uint32_t)(xs < 0.0f));
        const int32_t yi = (int32_t)((uint32_t)cvtt(ys) - (uint32_t)(ys < 0.0f));
        const bool mask = xi >= 0 && xi < (int32_t)sw && yi >= 0 && yi < (int32_t)sh;
        const bool mask1 = xi >= -1 && xi < (int32_t)sw && yi >= -1 && yi < (int32_t)sh;
        const float fx = xs - (float)xi, fy = ys - (float)yi;
        int32_t off = (int32_t)((uint32_t)yi * srcStride + (uint32_t)xi);
        off = max(min(off, limit), 0);
        const float p0 = mask ? (float)ldg8(src, n, off) : bf;
        const float p1 = mask1 ? (float)ldg8(src, n, (int64_t)off + 1) : bf;
        const float p2 = mask ? (float)ldg8(src, n, (int64_t)off + srcStride) : bf;
        const float p3 = mask1 ? (float)ldg8(src, n, (int64_t)off + srcStride + 1) : bf;
        const float omx = 1.0f - fx, omy = 1.0f - fy;
        float v = (p0 * omx) * omy;
        v = v + (p1 * fx) * omy;
        v = v + (p2 * omx) * fy;
        v = v + (p3 * fx) * fy;
        word |= sat8(cvtt(v)) << (8 * k);
    }
    *reinterpret_cast<uint32_t *>(dst + (uint64_t)y * dstStride + x0) = word;
}

// ---------------------------------------------------------------- LDS-staged bilinear warps --------
// The same per-pixel arithmetic as k_warp_affine / k_warp_perspective, for 64x16 output tiles whose
// source footprint is staged in LDS first: the tile's four corner pixels are mapped, their bounding
// box (plus a margin for the per-pixel rounding) is read row by row with dword loads at the linear
// byte addresses the reference reads -- a row's columns below 0 or past the stride are the bytes of
// the neighbouring rows, offsets outside the plane are the zero guard -- and each pixel then takes
// its four samples from LDS.  The staged box is only a cache: a pixel whose samples fall outside it
// (float rounding, a perspective tile that crosses the horizon, a box too large for LDS) or whose
// offset the reference clamps reads the plane directly, as the unstaged kernels do.  Needs a 4-byte
// aligned source with a stride that is a multiple of 4 (then every dword is wholly inside or wholly
// outside the plane).
constexpr int WTW = 64, WTH = 16;            // output tile: 16 threads x 4 pixels per row, 16 rows
constexpr int WLDS_DW = 1536;                // staged dwords per tile (6 KB: a 64x16 tile of a 2x zoom-out fits)

struct WarpBox { int r0, r1, cA, nd; bool on, inside; };

// bounding box of the source coordinates xs in [xlo, xhi], ys in [ylo, yhi]: rows and columns that a
// sample at trunc/floor of those coordinates and its +1 neighbours can touch, with two pixels of margin.
// inside: the box lies within the source image (columns [0, sw-1], rows [0, sh-1]); every pixel's
// samples are then inside the image, inside the box, and at offsets the reference never clamps, and
// its coordinates are small positive floats -- the conditions of the unchecked per-pixel path.
MVX_DEV WarpBox warp_box(float xlo, float xhi, float ylo, float yhi, uint32_t sw, uint32_t sh)
{
    WarpBox b{0, -1, 0, 0, false, false};
    if (!(xlo >= -65536.0f && xhi <= 65536.0f && ylo >= -65536.0f && yhi <= 65536.0f)) return b;
    const int c0 = (int)floorf(xlo) - 2, c1 = (int)floorf(xhi) + 3;
    b.r0 = (int)floorf(ylo) - 2;
    b.r1 = (int)floorf(yhi) + 3;
    b.cA = c0 & ~3;                                  // floor to a multiple of 4 (two's complement)
    b.nd = ((c1 - b.cA) >> 2) + 1;
    b.on = (int64_t)(b.r1 - b.r0 + 1) * b.nd <= WLDS_DW;    // rows and dwords are each below 2^18 here
    b.inside = b.on && c0 >= 0 && b.r0 >= 0 && c1 <= (int)sw - 1 && b.r1 <= (int)sh - 1;
    return b;
}

// stage rows r0..r1, columns cA..cA+4nd-1 (linear-address coordinates) into lds, row pitch 4nd bytes
MVX_DEV void warp_stage(uint32_t *lds, const WarpBox &b, const uint8_t *src, uint32_t stride, uint64_t n)
{
    const int rows = b.r1 - b.r0 + 1;
    for (int rr = (int)threadIdx.y; rr < rows; rr += (int)blockDim.y) {
        const int64_t base = (int64_t)(b.r0 + rr) * stride + b.cA;
        for (int k = (int)threadIdx.x; k < b.nd; k += (int)blockDim.x) {
            const int64_t lin = base + 4 * k;
            lds[rr * b.nd + k] = (lin >= 0 && (uint64_t)lin + 4 <= n) ? *reinterpret_cast<const uint32_t *>(src + lin) : 0u;
        }
    }
}

// the bytes at LDS offsets o and o+1 (o >= 0) in bits 0..15, from two dword-aligned reads and a byte
// shift: byte pairs read as 16-bit loads at odd offsets are misaligned, which the LDS serves far more
// slowly.  The stage arrays hold one dword past the box, so the second read stays inside them.
MVX_DEV uint32_t lds_pair(const uint32_t *lds, int o)
{
    const uint32_t lo = lds[o >> 2], hi = lds[(o >> 2) + 1];
    return __builtin_amdgcn_alignbyte(hi, lo, (uint32_t)o & 3u);
}

// the four samples at (yi, xi), (yi, xi+1), (yi+1, xi), (yi+1, xi+1) when the box holds them
MVX_DEV bool warp_lds4(const uint32_t *lds, const WarpBox &b, int32_t xi, int32_t yi, float &p0, float &p1, float &p2, float &p3)
{
    const int pitch = 4 * b.nd;
    if (!b.on || yi < b.r0 || yi >= b.r1 || xi < b.cA || xi >= b.cA + pitch - 1) return false;
    const int o = (yi - b.r0) * pitch + (xi - b.cA);
    const uint32_t t = lds_pair(lds, o), u = lds_pair(lds, o + pitch);
    p0 = (float)(t & 0xffu);
    p1 = (float)((t >> 8) & 0xffu);
    p2 = (float)(u & 0xffu);
    p3 = (float)((u >> 8) & 0xffu);
    return true;
}

template <bool BORDER>
__global__ void __launch_bounds__(256)
k_warp_affine_lds(uint32_t dw, uint32_t dh, uint8_t *__restrict__ dst, uint32_t dstStride, uint32_t sw, uint32_t sh,
                  const uint8_t *__restrict__ src, uint32_t srcStride, Affine m, int32_t limit, uint32_t border)
{
    __shared__ uint32_t stage[WLDS_DW + 1];
    const uint64_t n = (uint64_t)sh * srcStride;
    const float fw = (float)sw, fh = (float)sh;
    // tile corners (the last thread of a row stores pixels up to x0 + 63)
    const uint32_t tx0 = blockIdx.x * WTW, ty0 = blockIdx.y * WTH;
    WarpBox b;
    {
        const float xa = (float)tx0, xb = (float)(tx0 + WTW - 1), ya = (float)ty0, yb = (float)(ty0 + WTH - 1);
        const float sx[4] = {m.r00 * xa + (ya * m.r01 + m.c1), m.r00 * xb + (ya * m.r01 + m.c1), m.r00 * xa + (yb * m.r01 + m.c1),
                             m.r00 * xb + (yb * m.r01 + m.c1)};
        const float sy[4] = {m.r10 * xa + (ya * m.r11 + m.c2), m.r10 * xb + (ya * m.r11 + m.c2), m.r10 * xa + (yb * m.r11 + m.c2),
                             m.r10 * xb + (yb * m.r11 + m.c2)};
        b = warp_box(fminf(fminf(sx[0], sx[1]), fminf(sx[2], sx[3])), fmaxf(fmaxf(sx[0], sx[1]), fmaxf(sx[2], sx[3])),
                     fminf(fminf(sy[0], sy[1]), fminf(sy[2], sy[3])), fmaxf(fmaxf(sy[0], sy[1]), fmaxf(sy[2], sy[3])), sw, sh);
    }
    if (b.on) warp_stage(stage, b, src, srcStride, n);
    __syncthreads();
    const uint32_t x0 = tx0 + threadIdx.x * 4;
    const uint32_t y = ty0 + threadIdx.y;
    if (x0 >= dw || y >= dh) return;
    const float xd = (float)y * m.r01 + m.c1, yd = (float)y * m.r11 + m.c2;
    const uint32_t *lds = stage;
    uint32_t word = 0;
    // (a tile holding padding pixels -- computed from column 0's coefficients -- takes the checked path)
    if (b.inside && tx0 + WTW <= dw) {
        // every sample in the image and the box: no border test, clamp, range check or fallback
        const int pitch = 4 * b.nd;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            // every column of the tile is below dw here
            const float xs = m.r00 * (float)(x0 + k) + xd, ys = m.r10 * (float)(x0 + k) + yd;
            const int32_t xi = (int32_t)xs, yi = (int32_t)ys;
            const float fx = xs - (float)xi, fy = ys - (float)yi;
            const int o = (yi - b.r0) * pitch + (xi - b.cA);
            const uint32_t t = lds_pair(lds, o), u = lds_pair(lds, o + pitch);
            const float p0 = (float)(t & 0xffu), p1 = (float)((t >> 8) & 0xffu), p2 = (float)(u & 0xffu),
                        p3 = (float)((u >> 8) & 0xffu);
            const float omx = 1.0f - fx, omy = 1.0f - fy;
            const float a = (p0 * omx) * omy, bb = (p1 * fx) * omy, c = (p2 * omx) * fy, d = (p3 * fx) * fy;
            word |= sat8(__float2int_rn((a + bb) + (c + d))) << (8 * k);
        }
        *reinterpret_cast<uint32_t *>(dst + (uint64_t)y * dstStride + x0) = word;
        return;
    }
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const bool col = x0 + k < dw;
        const float xs = (col ? m.r00 * (float)(x0 + k) : 0.0f) + xd, ys = (col ? m.r10 * (float)(x0 + k) : 0.0f) + yd;
        const int32_t xi = cvtt(xs), yi = cvtt(ys);
        const float fx = xs - (float)xi, fy = ys - (float)yi;
        const int32_t raw = (int32_t)((uint32_t)yi * srcStride + (uint32_t)xi);
        const int32_t off = BORDER ? max(min(raw, limit), 0) : raw;
        float p0, p1, p2, p3;
        if (off != raw || !warp_lds4(lds, b, xi, yi, p0, p1, p2, p3)) {
            p0 = (float)ldg8(src, n, off);
            p1 = (float)ldg8(src, n, (int64_t)off + 1);
            p2 = (float)ldg8(src, n, (int64_t)off + srcStride);
            p3 = (float)ldg8(src, n, (int64_t)off + srcStride + 1);
        }
        const float omx = 1.0f - fx, omy = 1.0f - fy;
        const float a = (p0 * omx) * omy, bb = (p1 * fx) * omy, c = (p2 * omx) * fy, d = (p3 * fx) * fy;
        const float v = (a + bb) + (c + d);
        uint32_t px = sat8(cvtn(v));
        if (BORDER && !(xs >= 0.0f && xs < fw && ys >= 0.0f && ys < fh)) px = border;
        word |= px << (8 * k);
    }
    *reinterpret_cast<uint32_t *>(dst + (uint64_t)y * dstStride + x0) = word;
}

// the perspective tile is staged only when the four corner denominators share a sign (the tile then
// maps to a convex quadrilateral inside the corners' bounding box)
__global__ void __launch_bounds__(256)
k_warp_perspective_lds(uint32_t dw, uint32_t dh, uint8_t *__restrict__ dst, uint32_t dstStride, uint32_t sw, uint32_t sh,
                       const uint8_t *__restrict__ src, uint32_t srcStride, Persp m, uint32_t border)
{
    __shared__ uint32_t stage[WLDS_DW + 1];
    const uint64_t n = (uint64_t)sh * srcStride;
    const int32_t limit = (int32_t)(sh * srcStride - 1u);
    const uint32_t tx0 = blockIdx.x * WTW, ty0 = blockIdx.y * WTH;
    WarpBox b{0, -1, 0, 0, false};
    {
        float sx[4], sy[4], sz[4];
#pragma unroll
        for (int c = 0; c < 4; c++) {
            const float xf = (float)(tx0 + ((c & 1) ? WTW - 1 : 0)), yf = (float)(ty0 + ((c & 2) ? WTH - 1 : 0));
            sz[c] = m.g * xf + (yf * m.h + m.i);
            const float z = 1.0f / sz[c];
            sx[c] = (m.a * xf + (yf * m.b + m.c)) * z;
            sy[c] = (m.d * xf + (yf * m.e + m.f)) * z;
        }
        const bool pos = sz[0] > 0.0f && sz[1] > 0.0f && sz[2] > 0.0f && sz[3] > 0.0f;
        const bool neg = sz[0] < 0.0f && sz[1] < 0.0f && sz[2] < 0.0f && sz[3] < 0.0f;
        if (pos || neg)
            b = warp_box(fminf(fminf(sx[0], sx[1]), fminf(sx[2], sx[3])), fmaxf(fmaxf(sx[0], sx[1]), fmaxf(sx[2], sx[3])),
                         fminf(fminf(sy[0], sy[1]), fminf(sy[2], sy[3])), fmaxf(fmaxf(sy[0], sy[1]), fmaxf(sy[2], sy[3])), sw, sh);
    }
    if (b.on) warp_stage(stage, b, src, srcStride, n);
    __syncthreads();
    const uint32_t x0 = tx0 + threadIdx.x * 4;
    const uint32_t y = ty0 + threadIdx.y;
    if (x0 >= dw || y >= dh) return;
    const float xd = (float)y * m.b + m.c, yd = (float)y * m.e + m.f, zd = (float)y * m.h + m.i;
    const float bf = (float)border;
    const uint32_t *lds = stage;
    uint32_t word = 0;
    // (a tile holding padding pixels -- computed from column 0's coefficients -- takes the checked path)
    if (b.inside && tx0 + WTW <= dw) {
        // every sample in the image and the box (coordinates positive: floor is truncation)
        const int pitch = 4 * b.nd;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            // every column of the tile is below dw here
            const float xf = (float)(x0 + k);
            const float z = 1.0f / (m.g * xf + zd);
            const float xs = (m.a * xf + xd) * z, ys = (m.d * xf + yd) * z;
            const int32_t xi = (int32_t)xs, yi = (int32_t)ys;
            const float fx = xs - (float)xi, fy = ys - (float)yi;
            const int o = (yi - b.r0) * pitch + (xi - b.cA);
            const uint32_t t = lds_pair(lds, o), u = lds_pair(lds, o + pitch);
            const float p0 = (float)(t & 0xffu), p1 = (float)((t >> 8) & 0xffu), p2 = (float)(u & 0xffu),
                        p3 = (float)((u >> 8) & 0xffu);
            const float omx = 1.0f - fx, omy = 1.0f - fy;
            float v = (p0 * omx) * omy;
            v = v + (p1 * fx) * omy;
            v = v + (p2 * omx) * fy;
            v = v + (p3 * fx) * fy;
            word |= sat8((int32_t)v) << (8 * k);
        }
        *reinterpret_cast<uint32_t *>(dst + (uint64_t)y * dstStride + x0) = word;
        return;
    }
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const bool col = x0 + k < dw;
        const float xf = (float)(x0 + k);
        const float z = 1.0f / ((col ? m.g * xf : 0.0f) + zd);
        const float xs = ((col ? m.a * xf : 0.0f) + xd) * z, ys = ((col ? m.d * xf : 0.0f) + yd) * z;
        const int32_t xi = (int32_t)((uint32_t)cvtt(xs) - (uint32_t)(xs < 0.0f));
        const int32_t yi = (int32_t)((uint32_t)cvtt(ys) - (uint32_t)(ys < 0.0f));
        const bool mask = xi >= 0 && xi < (int32_t)sw && yi >= 0 && yi < (int32_t)sh;
        const bool mask1 = xi >= -1 && xi < (int32_t)sw && yi >= -1 && yi < (int32_t)sh;
        const float fx = xs - (float)xi, fy = ys - (float)yi;
        const int32_t raw = (int32_t)((uint32_t)yi * srcStride + (uint32_t)xi);
        const int32_t off = max(min(raw, limit), 0);
        float q0, q1, q2, q3;
        if (off != raw || !warp_lds4(lds, b, xi, yi, q0, q1, q2, q3)) {
            q0 = mask ? (float)ldg8(src, n, off) : 0.0f;
            q1 = mask1 ? (float)ldg8(src, n, (int64_t)off + 1) : 0.0f;
            q2 = mask ? (float)ldg8(src, n, (int64_t)off + srcStride) : 0.0f;
            q3 = mask1 ? (float)ldg8(src, n, (int64_t)off + srcStride + 1) : 0.0f;
        }
        const float p0 = mask ? q0 : bf, p1 = mask1 ? q1 : bf, p2 = mask ? q2 : bf, p3 = mask1 ? q3 : bf;
        const float omx = 1.0f - fx, omy = 1.0f - fy;
        float v = (p0 * omx) * omy;
        v = v + (p1 * fx) * omy;
        v = v + (p2 * omx) * fy;
        v = v + (p3 * fx) * fy;
        word |= sat8(cvtt(v)) << (8 * k);
    }
    *reinterpret_cast<uint32_t *>(dst + (uint64_t)y * dstStride + x0) = word;
}

// HafCpu_Remap_U8_U8_Bilinear[_Constant] (ago_haf_cpu_geometric.cpp:249-504).  The map is read at
// byte y*mapStride + 4x for every written pixel (pixels past w use the following entries, as the
// reference's 16-byte map loads do); entries past the end of the map read as (0, 0), the zeroed
// tail of the reference's map allocation.
template <bool CONSTANT>
__global__ void __launch_bounds__(256)
k_remap(uint32_t dw, uint32_t dh, uint8_t *__restrict__ dst, uint32_t dstStride, uint32_t sh,
        const uint8_t *__restrict__ src, uint32_t srcStride, const uint8_t *__restrict__ map, uint32_t mapStride,
        uint64_t mapBytes, uint32_t border)
{
    const uint32_t x0 = (blockIdx.x * blockDim.x + threadIdx.x) * 4;
    const uint32_t y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x0 >= dw || y >= dh) return;
    const uint64_t n = (uint64_t)sh * srcStride;
    uint32_t word = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const uint64_t mo = (uint64_t)y * mapStride + 4ull * (x0 + k);
        const uint32_t e = (mo + 4 <= mapBytes) ? *reinterpret_cast<const uint32_t *>(map + mo) : 0u;
        uint32_t mx = e & 0xffffu, my = e >> 16;
        const bool xbad = mx == 0xffffu;
        if (mx == 0xffffu) mx = CONSTANT ? 0u : 8u;
        if (my == 0xffffu) my = CONSTANT ? 0u : 8u;
        const int32_t fx = (int32_t)(mx & 7u), fy = (int32_t)(my & 7u);
        const int64_t o = (int64_t)(my >> 3) * srcStride + (int64_t)(mx >> 3);
        const int32_t p00 = (int32_t)ldg8(src, n, o), p10 = (int32_t)ldg8(src, n, o + 1);
        const int32_t p01 = (int32_t)ldg8(src, n, o + srcStride), p11 = (int32_t)ldg8(src, n, o + srcStride + 1);
        int32_t v = (p00 * (8 - fx) * (8 - fy) + p10 * fx * (8 - fy) + p01 * (8 - fx) * fy + p11 * fx * fy + 32) >> 6;
        if (CONSTANT && xbad) v = (int32_t)border;
        word |= sat8(v) << (8 * k);
    }
    *reinterpret_cast<uint32_t *>(dst + (uint64_t)y * dstStride + x0) = word;
}

// HafCpu_WarpAffine_U8_U8_Nearest[_Constant] (ago_haf_cpu_geometric.cpp:508-823): trunc(xs), trunc(ys),
// offset yi*stride + xi.  BORDER (the corner test failed): mask 0 <= xs < W, 0 <= ys < H and inside
// the valid rectangle [vl, vr) x [vt, vb), offset clamped to [0, H*stride - 1], the border value
// outside the mask (0 for UNDEFINED); otherwise the plain read.
template <bool BORDER>
__global__ void __launch_bounds__(256)
k_warp_affine_nn(uint32_t dw, uint32_t dh, uint8_t *__restrict__ dst, uint32_t dstStride, uint32_t sw, uint32_t sh,
                 const uint8_t *__restrict__ src, uint32_t srcStride, Affine m, uint32_t border, float vl, float vt, float vr,
                 float vb)
{
    const uint32_t x0 = (blockIdx.x * blockDim.x + threadIdx.x) * 4;
    const uint32_t y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x0 >= dw || y >= dh) return;
    const uint64_t n = (uint64_t)sh * srcStride;
    const int32_t limit = (int32_t)(sh * srcStride - 1u);
    const float fw = (float)sw, fh = (float)sh;
    const float xd = (float)y * m.r01 + m.c1, yd = (float)y * m.r11 + m.c2;
    uint32_t word = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const bool col = x0 + k < dw;
        const float xs = (col ? m.r00 * (float)(x0 + k) : 0.0f) + xd, ys = (col ? m.r10 * (float)(x0 + k) : 0.0f) + yd;
        const int32_t xi = cvtt(xs), yi = cvtt(ys);
        int32_t off = (int32_t)((uint32_t)yi * srcStride + (uint32_t)xi);
        uint32_t px;
        if (BORDER) {
            const bool mask = xs >= 0.0f && xs < fw && ys >= 0.0f && ys < fh && xs >= vl && xs < vr && ys >= vt && ys < vb;
            off = max(min(off, limit), 0);
            px = mask ? ldg8(src, n, off) : border;
        } else {
            px = ldg8(src, n, off);
        }
        word |= px << (8 * k);
    }
    *reinterpret_cast<uint32_t *>(dst + (uint64_t)y * dstStride + x0) = word;
}

// HafCpu_WarpPerspective_U8_U8_Nearest_Constant (ago_haf_cpu_geometric.cpp:1406-1558; UNDEFINED is
// border 0, :1372-1385): always the masked loop, mask 0 <= xs < W, 0 <= ys < H, offset
// trunc(ys)*stride + trunc(xs) clamped to [0, H*stride - 1].
__global__ void __launch_bounds__(256)
k_warp_perspective_nn(uint32_t dw, uint32_t dh, uint8_t *__restrict__ dst, uint32_t dstStride, uint32_t sw, uint32_t sh,
                      const uint8_t *__restrict__ src, uint32_t srcStride, Persp m, uint32_t border)
{
    const uint32_t x0 = (blockIdx.x * blockDim.x + threadIdx.x) * 4;
    const uint32_t y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x0 >= dw || y >= dh) return;
    const uint64_t n = (uint64_t)sh * srcStride;
    const int32_t limit = (int32_t)(sh * srcStride - 1u);
    const float fw = (float)sw, fh = (float)sh;
    const float xd = (float)y * m.b + m.c, yd = (float)y * m.e + m.f, zd = (float)y * m.h + m.i;
    uint32_t word = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const bool col = x0 + k < dw;
        const float xf = (float)(x0 + k);
        const float z = 1.0f / ((col ? m.g * xf : 0.0f) + zd);
        const float xs = ((col ? m.a * xf : 0.0f) + xd) * z, ys = ((col ? m.d * xf : 0.0f) + yd) * z;
        const bool mask = xs >= 0.0f && xs < fw && ys >= 0.0f && ys < fh;
        int32_t off = (int32_t)((uint32_t)cvtt(ys) * srcStride + (uint32_t)cvtt(xs));
        off = max(min(off, limit), 0);
        word |= (mask ? ldg8(src, n, off) : border) << (8 * k);
    }
    *reinterpret_cast<uint32_t *>(dst + (uint64_t)y * dstStride + x0) = word;
}

// HafCpu_Remap_U8_U8_Nearest[_Constant] (ago_haf_cpu_geometric.cpp:72-247).  Columns below w & ~3 (the
// reference's 4-pixel body), per 16-bit lane v: coordinate (v >> 3) + ((v & 7) > 3); a lane of 0xffff
// reads as 7 (UNDEFINED: coordinate 1) or 0 (CONSTANT, where the pixel is then the border value, as
// it is when the offset passes H*stride - 1).  The last w % 4 columns (its byte loop) take each lane
// as a signed 16-bit s: (s >> 3) + ((s & 7) >> 2), and s == -1 gives 0 (UNDEFINED) or the border value
// itself as the coordinate (CONSTANT); no bounds test there (offsets outside the plane read 0).
template <bool CONSTANT>
__global__ void __launch_bounds__(256)
k_remap_nn(uint32_t dw, uint32_t dh, uint8_t *__restrict__ dst, uint32_t dstStride, uint32_t sh,
           const uint8_t *__restrict__ src, uint32_t srcStride, const uint8_t *__restrict__ map, uint32_t mapStride,
           uint32_t border)
{
    const uint32_t x0 = (blockIdx.x * blockDim.x + threadIdx.x) * 4;
    const uint32_t y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x0 >= dw || y >= dh) return;
    const uint64_t n = (uint64_t)sh * srcStride;
    const int32_t limit = (int32_t)(sh * srcStride - 1u);
    const uint32_t *row = reinterpret_cast<const uint32_t *>(map + (uint64_t)y * mapStride);
    uint8_t *d = dst + (uint64_t)y * dstStride + x0;
    if (x0 + 4 <= dw) {
        uint32_t word = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const uint32_t e = row[x0 + k];
            uint32_t mx = e & 0xffffu, my = e >> 16;
            const bool bad = mx == 0xffffu || my == 0xffffu;
            if (mx == 0xffffu) mx = CONSTANT ? 0u : 7u;
            if (my == 0xffffu) my = CONSTANT ? 0u : 7u;
            const uint32_t xi = (mx >> 3) + ((mx & 7u) > 3u), yi = (my >> 3) + ((my & 7u) > 3u);
            const int32_t off = (int32_t)(yi * srcStride + xi);
            const uint32_t px = (CONSTANT && (bad || off > limit)) ? border : ldg8(src, n, off);
            word |= px << (8 * k);
        }
        *reinterpret_cast<uint32_t *>(d) = word;
    } else {
        for (uint32_t k = 0; k < dw - x0; k++) {
            const uint32_t e = row[x0 + k];
            const int32_t sx = (int16_t)(e & 0xffffu), sy = (int16_t)(e >> 16);
            const int32_t xi = sx != -1 ? (sx >> 3) + ((sx & 7) >> 2) : (CONSTANT ? (int32_t)border : 0);
            const int32_t yi = sy != -1 ? (sy >> 3) + ((sy & 7) >> 2) : (CONSTANT ? (int32_t)border : 0);
            const uint32_t off = (uint32_t)yi * srcStride + (uint32_t)xi;
            d[k] = (uint8_t)ldg8(src, n, (int64_t)off);
        }
    }
}

// RGB remap for the stitch path: the reference registers REMAP_U24_U24_BILINEAR but executes nothing
// (ago_kernel_api.cpp:18535-18539), so each channel is defined as the U8 bilinear remap of that channel
// (HafCpu_Remap_U8_U8_Bilinear[_Constant], ago_haf_cpu_geometric.cpp:249-504) over the same 13.3 table:
// every read of a valid entry stays inside the image (the table marks x >= W-1 and y >= H-1 invalid),
// so packed and planar layouts give the same bytes.  Exactly w pixels per row are written.
// one packed-RGB output pixel of the bilinear remap from its 13.3 table entry `e`
template <bool CONSTANT>
MVX_DEV void remap_u24_px(const uint8_t *__restrict__ src, uint32_t srcStride, uint32_t sw, uint32_t sh, uint32_t e, uint32_t border,
                          int32_t (&v)[3])
{
    uint32_t mx = e & 0xffffu, my = e >> 16;
    const bool xbad = mx == 0xffffu;
    if (mx == 0xffffu) mx = CONSTANT ? 0u : 8u;
    if (my == 0xffffu) my = CONSTANT ? 0u : 8u;
    const int32_t fx = (int32_t)(mx & 7u), fy = (int32_t)(my & 7u);
    const uint32_t xi = min(mx >> 3, sw - 1), yi = min(my >> 3, sh - 1);
    const uint32_t x1 = min(xi + 1, sw - 1), y1 = min(yi + 1, sh - 1);
    const uint8_t *r0 = src + (uint64_t)yi * srcStride, *r1 = src + (uint64_t)y1 * srcStride;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const int32_t p00 = r0[3 * xi + c], p10 = r0[3 * x1 + c], p01 = r1[3 * xi + c], p11 = r1[3 * x1 + c];
        int32_t t = (p00 * (8 - fx) * (8 - fy) + p10 * fx * (8 - fy) + p01 * (8 - fx) * fy + p11 * fx * fy + 32) >> 6;
        if (CONSTANT && xbad) t = (int32_t)border;
        v[c] = (int32_t)sat8(t);
    }
}

template <bool CONSTANT>
__global__ void __launch_bounds__(256)
k_remap_u24(uint32_t dw, uint32_t dh, uint8_t *__restrict__ dst, uint32_t dstStride, uint32_t sw, uint32_t sh,
            const uint8_t *__restrict__ src, uint32_t srcStride, const uint8_t *__restrict__ map, uint32_t mapStride, uint32_t border)
{
    const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= dw || y >= dh) return;
    const uint32_t e = *reinterpret_cast<const uint32_t *>(map + (uint64_t)y * mapStride + 4ull * x);
    int32_t v[3];
    remap_u24_px<CONSTANT>(src, srcStride, sw, sh, e, border, v);
    uint8_t *d = dst + (uint64_t)y * dstStride + 3ull * x;
#pragma unroll
    for (int c = 0; c < 3; c++) d[c] = (uint8_t)v[c];
}

// Stitch merge: the weighted sum of the remapped camera images (the role of Loom's merge kernel,
// amd_loomsl/kernels/merge.cpp:223-290, which scales U8 weights by 1/255 and accumulates the cameras in
// order; Loom is OpenCL-only and has no CPU form).  Per channel acc = sum over k in order of
// ((float)w_k * (1/255)) * (float)p_k, each product and sum rounded separately; out = acc rounded to
// nearest (ties to even), saturated to U8.
struct MergeArgs {
    const uint8_t *img[MVX_STITCH_MAX_CAMERAS];
    const uint8_t *wt[MVX_STITCH_MAX_CAMERAS];
    uint32_t imgStride[MVX_STITCH_MAX_CAMERAS], wtStride[MVX_STITCH_MAX_CAMERAS];
    uint32_t count;
};

__global__ void __launch_bounds__(256)
k_stitch_merge(uint32_t w, uint32_t h, uint8_t *__restrict__ dst, uint32_t dstStride, MergeArgs a)
{
    const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= w || y >= h) return;
    const float k255 = 1.0f / 255.0f;
    float acc[3] = {0.0f, 0.0f, 0.0f};
    for (uint32_t k = 0; k < a.count; k++) {
        const float wk = (float)a.wt[k][(uint64_t)y * a.wtStride[k] + x] * k255;
        const uint8_t *p = a.img[k] + (uint64_t)y * a.imgStride[k] + 3ull * x;
#pragma unroll
        for (int c = 0; c < 3; c++) acc[c] = acc[c] + wk * (float)p[c];
    }
    uint8_t *d = dst + (uint64_t)y * dstStride + 3ull * x;
#pragma unroll
    for (int c = 0; c < 3; c++) d[c] = (uint8_t)min(max(__float2int_rn(acc[c]), 0), 255);
}

// Loom's merge with its camera-selection maps (amd_loomsl/kernels/merge.cpp:176-385, generated OpenCL;
// Loom has no CPU form).  Per group of 8 output pixels of a row, camId0 (U8) selects:
//   < 31   one camera, copied (its pixel unweighted);
//   == 31  no camera: the output pixels are not written;
//   > 31   a blend of the cameras whose 5-bit ids camId1 (U16: bits 0-4, 5-9, 10-14) and camId2 (U16, the
//          same) list: the first two of camId1 always, its third when camId0 > 128, then camId2's first,
//          second and third when camId0 > 129, > 130, > 131; an id of 31 contributes nothing.
// A blended camera contributes w * p with w = (float)weight * 0.003922f -- the kernel source prints its
// factor 1/255 with "%f" (merge.cpp:226,256), so that is the constant the kernel runs with -- summed in
// list order, each product and sum rounded (OpenCL's mad taken unfused).  The sum is packed to U8 with
// round-to-nearest-even and saturation (amd_pack, v_cvt_pk_u8_f32).  Images and weights of camera c are
// rows c*height .. c*height + height-1 of the stacked planes.  A camera id at or above `count` (no such
// camera) is treated as 31 here instead of reading past the stack.
struct LoomArgs {
    const uint8_t *camId0;
    const uint16_t *camId1, *camId2;
    const uint8_t *wt;
    uint32_t camId0Stride, camId1Stride, camId2Stride, wtStride;
    uint32_t width, height, count;
};

constexpr float LOOM_WEIGHT_FACTOR = 0.003922f;

MVX_DEV uint32_t loom_pack(float v) { return (uint32_t)min(max(__float2int_rn(v), 0), 255); }

// Src(c, x, y, rgb) fetches camera c's pixel (x, y) of the equirectangular frame
template <class Src>
MVX_DEV bool loom_merge_px(const LoomArgs &a, uint32_t x, uint32_t y, const Src &src, uint32_t (&out)[3])
{
    const uint32_t g = x >> 3;
    const uint32_t sel = a.camId0[(uint64_t)y * a.camId0Stride + g];
    if (sel == 31u) return false;
    int32_t p[3];
    if (sel < 31u) {
        if (sel >= a.count) return false;
        src(sel, x, y, p);
        for (int c = 0; c < 3; c++) out[c] = (uint32_t)p[c];
        return true;
    }
    const uint32_t s1 = *(const uint16_t *)((const uint8_t *)a.camId1 + (uint64_t)y * a.camId1Stride + 2ull * g);
    const uint32_t s2 = sel > 129u ? *(const uint16_t *)((const uint8_t *)a.camId2 + (uint64_t)y * a.camId2Stride + 2ull * g) : 0u;
    const int n = sel > 131u ? 6 : sel > 130u ? 5 : sel > 129u ? 4 : sel > 128u ? 3 : 2;
    float fa[3] = {0.0f, 0.0f, 0.0f};
    for (int k = 0; k < n; k++) {
        const uint32_t id = ((k < 3 ? s1 : s2) >> (5 * (k % 3))) & 31u;
        float w = 0.0f;
        p[0] = p[1] = p[2] = 0;
        if (id < 31u && id < a.count) {
            src(id, x, y, p);
            w = (float)a.wt[((uint64_t)id * a.height + y) * a.wtStride + x] * LOOM_WEIGHT_FACTOR;
        }
#pragma unroll
        for (int c = 0; c < 3; c++) fa[c] = fa[c] + w * (float)p[c];
    }
    for (int c = 0; c < 3; c++) out[c] = loom_pack(fa[c]);
    return true;
}

struct StackedSrc {   // the remapped cameras, stacked RGB planes
    const uint8_t *img;
    uint32_t stride, height;
    MVX_DEV void operator()(uint32_t c, uint32_t x, uint32_t y, int32_t (&p)[3]) const
    {
        const uint8_t *q = img + ((uint64_t)c * height + y) * stride + 3ull * x;
        p[0] = q[0]; p[1] = q[1]; p[2] = q[2];
    }
};

__global__ void __launch_bounds__(256)
k_stitch_loom_merge(uint8_t *__restrict__ dst, uint32_t dstStride, LoomArgs a, StackedSrc s)
{
    const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= a.width || y >= a.height) return;
    uint32_t o[3];
    if (!loom_merge_px(a, x, y, s, o)) return;
    uint8_t *d = dst + (uint64_t)y * dstStride + 3ull * x;
    d[0] = (uint8_t)o[0]; d[1] = (uint8_t)o[1]; d[2] = (uint8_t)o[2];
}

struct RemapSrc {     // the cameras themselves, remapped on the fly (bilinear, constant border 0)
    const uint8_t *cam[MVX_STITCH_MAX_CAMERAS];
    uint32_t camStride[MVX_STITCH_MAX_CAMERAS];
    const uint8_t *map;      // stacked 13.3 tables, rows c*height ..
    uint32_t mapStride, height, sw, sh;
    MVX_DEV void operator()(uint32_t c, uint32_t x, uint32_t y, int32_t (&p)[3]) const
    {
        const uint32_t e = *reinterpret_cast<const uint32_t *>(map + ((uint64_t)c * height + y) * mapStride + 4ull * x);
        remap_u24_px<true>(cam[c], camStride[c], sw, sh, e, 0u, p);
    }
};

// Remap and merge in one pass: each output pixel samples only the cameras its selection lists, so the
// warped camera frames never exist in HBM.  Same bytes as k_remap_u24<CONSTANT> per camera followed by
// k_stitch_loom_merge.
__global__ void __launch_bounds__(256)
k_stitch_loom_fused(uint8_t *__restrict__ dst, uint32_t dstStride, LoomArgs a, RemapSrc s)
{
    const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= a.width || y >= a.height) return;
    uint32_t o[3];
    if (!loom_merge_px(a, x, y, s, o)) return;
    uint8_t *d = dst + (uint64_t)y * dstStride + 3ull * x;
    d[0] = (uint8_t)o[0]; d[1] = (uint8_t)o[1]; d[2] = (uint8_t)o[2];
}

// vxSetRemapPoint (vx_api.cpp:7865-7894) for every entry: x = (u16)(sx*8 + 0.5f), y likewise,
// (0xffff, 0xffff) when sx < 0, sy < 0, sx >= W-1 or sy >= H-1 (3 fractional bits: sources below
// 8192 pixels, ago_util.cpp:2450-2459).
__global__ void __launch_bounds__(256)
k_remap_fixed(uint32_t dw, uint32_t dh, uint8_t *__restrict__ dstMap, uint32_t dstStride, const uint8_t *__restrict__ srcMap,
              uint32_t srcStride, float limX, float limY)
{
    const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= dw || y >= dh) return;
    const float2 s = *reinterpret_cast<const float2 *>(srcMap + (uint64_t)y * srcStride + 8ull * x);
    uint32_t fx = (uint32_t)(uint16_t)(int32_t)(s.x * 8.0f + 0.5f), fy = (uint32_t)(uint16_t)(int32_t)(s.y * 8.0f + 0.5f);
    if (s.x < 0.0f || s.y < 0.0f || s.x >= limX || s.y >= limY) fx = fy = 0xffffu;
    *reinterpret_cast<uint32_t *>(dstMap + (uint64_t)y * dstStride + 4ull * x) = fx | (fy << 16);
}

namespace {

bool warp_args_ok(uint32_t dw, uint32_t dh, const uint8_t *dst, uint32_t dstStride, uint32_t sw, uint32_t sh,
                  const uint8_t *src, uint32_t srcStride)
{
    if (!dst || !src || !dw || !dh || !sw || !sh) return false;
    if (((uintptr_t)dst & 3u) || (dstStride & 3u) || dstStride < ((dw + 3u) & ~3u)) return false;
    if (srcStride < sw || (uint64_t)sh * srcStride >= (1ull << 31)) return false;
    return true;
}

// the staged kernels read the source as dwords: a 4-byte aligned base and stride
bool staged_ok(const uint8_t *src, uint32_t srcStride) { return ((uintptr_t)src & 3u) == 0 && (srcStride & 3u) == 0; }

dim3 warp_grid(uint32_t dw, uint32_t dh, dim3 block) { return dim3((((dw + 3) / 4) + block.x - 1) / block.x, (dh + block.y - 1) / block.y); }

// the corner test of ago_haf_cpu_geometric.cpp:1142-1155 (host float arithmetic, no contraction)
bool affine_needs_border(uint32_t dw, uint32_t dh, uint32_t sw, uint32_t sh, const Affine &m)
{
    const float fw = (float)sw, fh = (float)sh, w = (float)dw, h = (float)dh;
    bool b = (m.c1 < 0) | (m.c2 < 0) | (m.c1 >= fw) | (m.c2 >= fh);
    float x1 = m.r00 * w + m.c1, y1 = m.r10 * w + m.c2;
    b |= (x1 < 0) | (y1 < 0) | (x1 >= fw) | (y1 >= fh);
    x1 = m.r01 * h + m.c1;
    y1 = m.r11 * h + m.c2;
    b |= (x1 < 0) | (y1 < 0) | (x1 >= fw) | (y1 >= fh);
    x1 = m.r00 * w + m.r01 * h + m.c1;
    y1 = m.r10 * w + m.r11 * h + m.c2;
    b |= (x1 < 0) | (y1 < 0) | (x1 >= fw) | (y1 >= fh);
    return b;
}

int warp_affine(mvx_stream_t stream, uint32_t dw, uint32_t dh, uint8_t *dst, uint32_t dstStride, uint32_t sw, uint32_t sh,
                const uint8_t *src, uint32_t srcStride, const float *matrix, bool constant, uint32_t border)
{
    if (!matrix || !warp_args_ok(dw, dh, dst, dstStride, sw, sh, src, srcStride)) return MVX_ERROR_INVALID_PARAMETERS;
    const Affine m{matrix[0], matrix[1], matrix[2], matrix[3], matrix[4], matrix[5]};
    const int32_t limit = constant ? (int32_t)(sh * srcStride - 1u) : (int32_t)((sh - 1u) * srcStride - 1u);
    if (staged_ok(src, srcStride)) {
        const dim3 block(WTW / 4, WTH), grid((dw + WTW - 1) / WTW, (dh + WTH - 1) / WTH);
        if (affine_needs_border(dw, dh, sw, sh, m))
            hipLaunchKernelGGL(k_warp_affine_lds<true>, grid, block, 0, stream_of(stream), dw, dh, dst, dstStride, sw, sh, src,
                               srcStride, m, limit, constant ? border : 0u);
        else
            hipLaunchKernelGGL(k_warp_affine_lds<false>, grid, block, 0, stream_of(stream), dw, dh, dst, dstStride, sw, sh, src,
                               srcStride, m, limit, 0u);
        MVX_LAUNCH_CHECK();
        return MVX_SUCCESS;
    }
    const dim3 block(64, 4);
    if (affine_needs_border(dw, dh, sw, sh, m))
        hipLaunchKernelGGL(k_warp_affine<true>, warp_grid(dw, dh, block), block, 0, stream_of(stream), dw, dh, dst, dstStride, sw,
                           sh, src, srcStride, m, limit, constant ? border : 0u);
    else
        hipLaunchKernelGGL(k_warp_affine<false>, warp_grid(dw, dh, block), block, 0, stream_of(stream), dw, dh, dst, dstStride, sw,
                           sh, src, srcStride, m, limit, 0u);
    MVX_LAUNCH_CHECK();
    return MVX_SUCCESS;
}

int warp_perspective(mvx_stream_t stream, uint32_t dw, uint32_t dh, uint8_t *dst, uint32_t dstStride, uint32_t sw, uint32_t sh,
                     const uint8_t *src, uint32_t srcStride, const float *matrix, uint32_t border)
{
    if (!matrix || !warp_args_ok(dw, dh, dst, dstStride, sw, sh, src, srcStride)) return MVX_ERROR_INVALID_PARAMETERS;
    const Persp m{matrix[0], matrix[1], matrix[2], matrix[3], matrix[4], matrix[5], matrix[6], matrix[7], matrix[8]};
    if (staged_ok(src, srcStride)) {
        const dim3 block(WTW / 4, WTH), grid((dw + WTW - 1) / WTW, (dh + WTH - 1) / WTH);
        hipLaunchKernelGGL(k_warp_perspective_lds, grid, block, 0, stream_of(stream), dw, dh, dst, dstStride, sw, sh, src, srcStride, m,
                           border);
        MVX_LAUNCH_CHECK();
        return MVX_SUCCESS;
    }
    const dim3 block(64, 4);
    hipLaunchKernelGGL(k_warp_perspective, warp_grid(dw, dh, block), block, 0, stream_of(stream), dw, dh, dst, dstStride, sw, sh, src,
                       srcStride, m, border);
    MVX_LAUNCH_CHECK();
    return MVX_SUCCESS;
}

int warp_affine_nn(mvx_stream_t stream, uint32_t dw, uint32_t dh, uint8_t *dst, uint32_t dstStride, uint32_t sw, uint32_t sh,
                   const uint8_t *src, uint32_t srcStride, const float *matrix, uint32_t border, const uint32_t *valid)
{
    if (!matrix || !warp_args_ok(dw, dh, dst, dstStride, sw, sh, src, srcStride)) return MVX_ERROR_INVALID_PARAMETERS;
    const Affine m{matrix[0], matrix[1], matrix[2], matrix[3], matrix[4], matrix[5]};
    const dim3 block(64, 4);
    const float vl = valid ? (float)valid[0] : 0.0f, vt = valid ? (float)valid[1] : 0.0f;
    const float vr = valid ? (float)valid[2] : (float)sw, vb = valid ? (float)valid[3] : (float)sh;
    if (affine_needs_border(dw, dh, sw, sh, m))
        hipLaunchKernelGGL(k_warp_affine_nn<true>, warp_grid(dw, dh, block), block, 0, stream_of(stream), dw, dh, dst, dstStride, sw,
                           sh, src, srcStride, m, border, vl, vt, vr, vb);
    else
        hipLaunchKernelGGL(k_warp_affine_nn<false>, warp_grid(dw, dh, block), block, 0, stream_of(stream), dw, dh, dst, dstStride, sw,
                           sh, src, srcStride, m, 0u, vl, vt, vr, vb);
    MVX_LAUNCH_CHECK();
    return MVX_SUCCESS;
}

int warp_perspective_nn(mvx_stream_t stream, uint32_t dw, uint32_t dh, uint8_t *dst, uint32_t dstStride, uint32_t sw, uint32_t sh,
                        const uint8_t *src, uint32_t srcStride, const float *matrix, uint32_t border)
{
    if (!matrix || !warp_args_ok(dw, dh, dst, dstStride, sw, sh, src, srcStride)) return MVX_ERROR_INVALID_PARAMETERS;
    const Persp m{matrix[0], matrix[1], matrix[2], matrix[3], matrix[4], matrix[5], matrix[6], matrix[7], matrix[8]};
    const dim3 block(64, 4);
    hipLaunchKernelGGL(k_warp_perspective_nn, warp_grid(dw, dh, block), block, 0, stream_of(stream), dw, dh, dst, dstStride, sw, sh,
                       src, srcStride, m, border);
    MVX_LAUNCH_CHECK();
    return MVX_SUCCESS;
}

int remap_nn(mvx_stream_t stream, uint32_t dw, uint32_t dh, uint8_t *dst, uint32_t dstStride, uint32_t sw, uint32_t sh,
             const uint8_t *src, uint32_t srcStride, const void *map, uint32_t mapStride, bool constant, uint32_t border)
{
    if (!map || ((uintptr_t)map & 3u) || (mapStride & 3u) || mapStride < 4u * dw ||
        !warp_args_ok(dw, dh, dst, dstStride, sw, sh, src, srcStride))
        return MVX_ERROR_INVALID_PARAMETERS;
    const dim3 block(64, 4);
    if (constant)
        hipLaunchKernelGGL(k_remap_nn<true>, warp_grid(dw, dh, block), block, 0, stream_of(stream), dw, dh, dst, dstStride, sh, src,
                           srcStride, (const uint8_t *)map, mapStride, border);
    else
        hipLaunchKernelGGL(k_remap_nn<false>, warp_grid(dw, dh, block), block, 0, stream_of(stream), dw, dh, dst, dstStride, sh, src,
                           srcStride, (const uint8_t *)map, mapStride, 0u);
    MVX_LAUNCH_CHECK();
    return MVX_SUCCESS;
}

int remap(mvx_stream_t stream, uint32_t dw, uint32_t dh, uint8_t *dst, uint32_t dstStride, uint32_t sw, uint32_t sh,
          const uint8_t *src, uint32_t srcStride, const void *map, uint32_t mapStride, bool constant, uint32_t border)
{
    if (!map || ((uintptr_t)map & 3u) || (mapStride & 3u) || mapStride < 4u * dw ||
        !warp_args_ok(dw, dh, dst, dstStride, sw, sh, src, srcStride))
        return MVX_ERROR_INVALID_PARAMETERS;
    const dim3 block(64, 4);
    const uint64_t mapBytes = (uint64_t)dh * mapStride;
    if (constant)
        hipLaunchKernelGGL(k_remap<true>, warp_grid(dw, dh, block), block, 0, stream_of(stream), dw, dh, dst, dstStride, sh, src,
                           srcStride, (const uint8_t *)map, mapStride, mapBytes, border);
    else
        hipLaunchKernelGGL(k_remap<false>, warp_grid(dw, dh, block), block, 0, stream_of(stream), dw, dh, dst, dstStride, sh, src,
                           srcStride, (const uint8_t *)map, mapStride, mapBytes, 0u);
    MVX_LAUNCH_CHECK();
    return MVX_SUCCESS;
}

}  // namespace
}  // namespace mvx

using namespace mvx;

extern "C" int mvx_WarpAffine_U8_U8_Bilinear(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight, uint8_t *pDst,
                                             uint32_t dstStride, uint32_t srcWidth, uint32_t srcHeight, const uint8_t *pSrc,
                                             uint32_t srcStride, const float *matrix)
{
    return warp_affine(stream, dstWidth, dstHeight, pDst, dstStride, srcWidth, srcHeight, pSrc, srcStride, matrix, false, 0);
}

extern "C" int mvx_WarpAffine_U8_U8_Bilinear_Constant(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight, uint8_t *pDst,
                                                      uint32_t dstStride, uint32_t srcWidth, uint32_t srcHeight,
                                                      const uint8_t *pSrc, uint32_t srcStride, const float *matrix,
                                                      uint8_t borderValue)
{
    return warp_affine(stream, dstWidth, dstHeight, pDst, dstStride, srcWidth, srcHeight, pSrc, srcStride, matrix, true, borderValue);
}

extern "C" int mvx_WarpPerspective_U8_U8_Bilinear(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight, uint8_t *pDst,
                                                  uint32_t dstStride, uint32_t srcWidth, uint32_t srcHeight, const uint8_t *pSrc,
                                                  uint32_t srcStride, const float *matrix)
{
    return warp_perspective(stream, dstWidth, dstHeight, pDst, dstStride, srcWidth, srcHeight, pSrc, srcStride, matrix, 0);
}

extern "C" int mvx_WarpPerspective_U8_U8_Bilinear_Constant(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight,
                                                           uint8_t *pDst, uint32_t dstStride, uint32_t srcWidth,
                                                           uint32_t srcHeight, const uint8_t *pSrc, uint32_t srcStride,
                                                           const float *matrix, uint8_t borderValue)
{
    return warp_perspective(stream, dstWidth, dstHeight, pDst, dstStride, srcWidth, srcHeight, pSrc, srcStride, matrix, borderValue);
}

// The reference's srcImageBufferSize (bytes of the source buffer) guards its kernels' reads; here the
// table only holds in-range points, so the size is checked once: it must cover the source plane.
static bool src_buffer_ok(uint32_t srcWidth, uint32_t srcHeight, uint32_t srcStride, uint32_t srcImageBufferSize)
{
    if (!srcWidth || !srcHeight) return true;
    return (uint64_t)(srcHeight - 1) * srcStride + srcWidth <= srcImageBufferSize;
}

extern "C" int mvx_Remap_U8_U8_Bilinear(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight, uint8_t *pDst,
                                        uint32_t dstStride, uint32_t srcWidth, uint32_t srcHeight, const uint8_t *pSrc,
                                        uint32_t srcStride, uint32_t srcImageBufferSize, const void *pMap, uint32_t mapStride)
{
    if (!src_buffer_ok(srcWidth, srcHeight, srcStride, srcImageBufferSize)) return MVX_ERROR_INVALID_PARAMETERS;
    return remap(stream, dstWidth, dstHeight, pDst, dstStride, srcWidth, srcHeight, pSrc, srcStride, pMap, mapStride, false, 0);
}

extern "C" int mvx_Remap_U8_U8_Bilinear_Constant(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight, uint8_t *pDst,
                                                 uint32_t dstStride, uint32_t srcWidth, uint32_t srcHeight, const uint8_t *pSrc,
                                                 uint32_t srcStride, const void *pMap, uint32_t mapStride, uint8_t borderValue)
{
    return remap(stream, dstWidth, dstHeight, pDst, dstStride, srcWidth, srcHeight, pSrc, srcStride, pMap, mapStride, true, borderValue);
}

extern "C" int mvx_WarpAffine_U8_U8_Nearest(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight, uint8_t *pDst,
                                            uint32_t dstStride, uint32_t srcWidth, uint32_t srcHeight, const uint8_t *pSrc,
                                            uint32_t srcStride, const float *matrix)
{
    return warp_affine_nn(stream, dstWidth, dstHeight, pDst, dstStride, srcWidth, srcHeight, pSrc, srcStride, matrix, 0, nullptr);
}

extern "C" int mvx_WarpAffine_U8_U8_Nearest_Constant(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight, uint8_t *pDst,
                                                     uint32_t dstStride, uint32_t srcWidth, uint32_t srcHeight,
                                                     const uint8_t *pSrc, uint32_t srcStride, const float *matrix,
                                                     uint8_t borderValue, const uint32_t *validRect)
{
    return warp_affine_nn(stream, dstWidth, dstHeight, pDst, dstStride, srcWidth, srcHeight, pSrc, srcStride, matrix, borderValue,
                          validRect);
}

extern "C" int mvx_WarpPerspective_U8_U8_Nearest(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight, uint8_t *pDst,
                                                 uint32_t dstStride, uint32_t srcWidth, uint32_t srcHeight, const uint8_t *pSrc,
                                                 uint32_t srcStride, const float *matrix)
{
    return warp_perspective_nn(stream, dstWidth, dstHeight, pDst, dstStride, srcWidth, srcHeight, pSrc, srcStride, matrix, 0);
}

extern "C" int mvx_WarpPerspective_U8_U8_Nearest_Constant(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight,
                                                          uint8_t *pDst, uint32_t dstStride, uint32_t srcWidth,
                                                          uint32_t srcHeight, const uint8_t *pSrc, uint32_t srcStride,
                                                          const float *matrix, uint8_t borderValue)
{
    return warp_perspective_nn(stream, dstWidth, dstHeight, pDst, dstStride, srcWidth, srcHeight, pSrc, srcStride, matrix,
                               borderValue);
}

extern "C" int mvx_Remap_U8_U8_Nearest(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight, uint8_t *pDst,
                                       uint32_t dstStride, uint32_t srcWidth, uint32_t srcHeight, const uint8_t *pSrc,
                                       uint32_t srcStride, uint32_t srcImageBufferSize, const void *pMap, uint32_t mapStride)
{
    if (!src_buffer_ok(srcWidth, srcHeight, srcStride, srcImageBufferSize)) return MVX_ERROR_INVALID_PARAMETERS;
    return remap_nn(stream, dstWidth, dstHeight, pDst, dstStride, srcWidth, srcHeight, pSrc, srcStride, pMap, mapStride, false, 0);
}

extern "C" int mvx_Remap_U8_U8_Nearest_Constant(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight, uint8_t *pDst,
                                                uint32_t dstStride, uint32_t srcWidth, uint32_t srcHeight, const uint8_t *pSrc,
                                                uint32_t srcStride, uint32_t srcImageBufferSize, const void *pMap,
                                                uint32_t mapStride, uint8_t borderValue)
{
    if (!src_buffer_ok(srcWidth, srcHeight, srcStride, srcImageBufferSize)) return MVX_ERROR_INVALID_PARAMETERS;
    return remap_nn(stream, dstWidth, dstHeight, pDst, dstStride, srcWidth, srcHeight, pSrc, srcStride, pMap, mapStride, true,
                    borderValue);
}

namespace {
int remap_u24(mvx_stream_t stream, uint32_t dw, uint32_t dh, uint8_t *dst, uint32_t dstStride, uint32_t sw, uint32_t sh, const uint8_t *src,
              uint32_t srcStride, const void *map, uint32_t mapStride, bool constant, uint32_t border)
{
    if (!dst || !src || !map || !dw || !dh || sw < 2 || sh < 2 || sw >= 8192 || sh >= 8192 || dstStride < 3 * dw || srcStride < 3 * sw ||
        ((uintptr_t)map & 3u) || (mapStride & 3u) || mapStride < 4u * dw)
        return MVX_ERROR_INVALID_PARAMETERS;
    const dim3 block(64, 4), grid((dw + 63) / 64, (dh + 3) / 4);
    if (constant)
        hipLaunchKernelGGL(k_remap_u24<true>, grid, block, 0, stream_of(stream), dw, dh, dst, dstStride, sw, sh, src, srcStride,
                           (const uint8_t *)map, mapStride, border);
    else
        hipLaunchKernelGGL(k_remap_u24<false>, grid, block, 0, stream_of(stream), dw, dh, dst, dstStride, sw, sh, src, srcStride,
                           (const uint8_t *)map, mapStride, 0u);
    MVX_LAUNCH_CHECK();
    return MVX_SUCCESS;
}
}  // namespace

extern "C" int mvx_Remap_U24_U24_Bilinear(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight, uint8_t *pDst, uint32_t dstStride,
                                          uint32_t srcWidth, uint32_t srcHeight, const uint8_t *pSrc, uint32_t srcStride, const void *pMap,
                                          uint32_t mapStride)
{
    return remap_u24(stream, dstWidth, dstHeight, pDst, dstStride, srcWidth, srcHeight, pSrc, srcStride, pMap, mapStride, false, 0);
}

extern "C" int mvx_Remap_U24_U24_Bilinear_Constant(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight, uint8_t *pDst,
                                                   uint32_t dstStride, uint32_t srcWidth, uint32_t srcHeight, const uint8_t *pSrc,
                                                   uint32_t srcStride, const void *pMap, uint32_t mapStride, uint8_t borderValue)
{
    return remap_u24(stream, dstWidth, dstHeight, pDst, dstStride, srcWidth, srcHeight, pSrc, srcStride, pMap, mapStride, true, borderValue);
}

extern "C" int mvx_StitchMerge_RGB(mvx_stream_t stream, uint32_t width, uint32_t height, uint8_t *pDst, uint32_t dstStride, uint32_t count,
                                  const uint8_t *const *pImages, const uint32_t *imageStrides, const uint8_t *const *pWeights,
                                  const uint32_t *weightStrides)
{
    if (!pDst || !pImages || !imageStrides || !pWeights || !weightStrides || count == 0 || count > MVX_STITCH_MAX_CAMERAS || !width ||
        !height || dstStride < 3 * width)
        return MVX_ERROR_INVALID_PARAMETERS;
    MergeArgs a{};
    a.count = count;
    for (uint32_t k = 0; k < count; k++) {
        if (!pImages[k] || !pWeights[k] || imageStrides[k] < 3 * width || weightStrides[k] < width) return MVX_ERROR_INVALID_PARAMETERS;
        a.img[k] = pImages[k];
        a.wt[k] = pWeights[k];
        a.imgStride[k] = imageStrides[k];
        a.wtStride[k] = weightStrides[k];
    }
    const dim3 block(64, 4), grid((width + 63) / 64, (height + 3) / 4);
    hipLaunchKernelGGL(k_stitch_merge, grid, block, 0, stream_of(stream), width, height, pDst, dstStride, a);
    MVX_LAUNCH_CHECK();
    return MVX_SUCCESS;
}

extern "C" int mvx_Remap_FixedMap_F32(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight, void *pFixedMap,
                                      uint32_t fixedStride, const float *pFloatMap, uint32_t floatStride, uint32_t srcWidth,
                                      uint32_t srcHeight)
{
    if (!pFixedMap || !pFloatMap || !dstWidth || !dstHeight || srcWidth < 2 || srcHeight < 2 || srcWidth >= 8192 || srcHeight >= 8192)
        return MVX_ERROR_INVALID_PARAMETERS;
    if (((uintptr_t)pFixedMap & 3u) || ((uintptr_t)pFloatMap & 7u) || (fixedStride & 3u) || (floatStride & 7u) ||
        fixedStride < 4u * dstWidth || floatStride < 8u * dstWidth)
        return MVX_ERROR_INVALID_PARAMETERS;
    const dim3 block(64, 4), grid((dstWidth + 63) / 64, (dstHeight + 3) / 4);
    hipLaunchKernelGGL(k_remap_fixed, grid, block, 0, stream_of(stream), dstWidth, dstHeight, (uint8_t *)pFixedMap, fixedStride,
                       (const uint8_t *)pFloatMap, floatStride, (float)(srcWidth - 1), (float)(srcHeight - 1));
    MVX_LAUNCH_CHECK();
    return MVX_SUCCESS;
}

namespace mvx {
static int loom_args(LoomArgs &a, uint32_t width, uint32_t height, uint32_t count, const uint8_t *pWeights, uint32_t weightStride,
                     const uint8_t *pCamId0, uint32_t camId0Stride, const uint16_t *pCamId1, uint32_t camId1Stride,
                     const uint16_t *pCamId2, uint32_t camId2Stride)
{
    const uint32_t groups = (width + 7) / 8;
    if (!width || !height || count == 0 || count > MVX_STITCH_MAX_CAMERAS || !pWeights || !pCamId0 || !pCamId1 || !pCamId2)
        return MVX_ERROR_INVALID_PARAMETERS;
    if (weightStride < width || camId0Stride < groups || camId1Stride < 2 * groups || camId2Stride < 2 * groups ||
        ((uintptr_t)pCamId1 & 1u) || ((uintptr_t)pCamId2 & 1u) || (camId1Stride & 1u) || (camId2Stride & 1u))
        return MVX_ERROR_INVALID_PARAMETERS;
    a = LoomArgs{pCamId0, pCamId1, pCamId2, pWeights, camId0Stride, camId1Stride, camId2Stride, weightStride, width, height, count};
    return MVX_SUCCESS;
}
}  // namespace mvx

extern "C" int mvx_StitchMerge_Loom_RGB(mvx_stream_t stream, uint32_t width, uint32_t height, uint8_t *pDst, uint32_t dstStride,
                                       uint32_t count, const uint8_t *pStacked, uint32_t stackedStride, const uint8_t *pWeights,
                                       uint32_t weightStride, const uint8_t *pCamId0, uint32_t camId0Stride, const uint16_t *pCamId1,
                                       uint32_t camId1Stride, const uint16_t *pCamId2, uint32_t camId2Stride)
{
    LoomArgs a;
    int rc = loom_args(a, width, height, count, pWeights, weightStride, pCamId0, camId0Stride, pCamId1, camId1Stride, pCamId2,
                       camId2Stride);
    if (rc) return rc;
    if (!pDst || !pStacked || dstStride < 3 * width || stackedStride < 3 * width) return MVX_ERROR_INVALID_PARAMETERS;
    const dim3 block(64, 4), grid((width + 63) / 64, (height + 3) / 4);
    hipLaunchKernelGGL(k_stitch_loom_merge, grid, block, 0, stream_of(stream), pDst, dstStride, a, StackedSrc{pStacked, stackedStride, height});
    MVX_LAUNCH_CHECK();
    return MVX_SUCCESS;
}

extern "C" int mvx_Stitch_Loom_RGB(mvx_stream_t stream, uint32_t width, uint32_t height, uint8_t *pDst, uint32_t dstStride,
                                  uint32_t count, const uint8_t *const *pCameras, const uint32_t *cameraStrides, uint32_t srcWidth,
                                  uint32_t srcHeight, const void *pMaps, uint32_t mapStride, const uint8_t *pWeights,
                                  uint32_t weightStride, const uint8_t *pCamId0, uint32_t camId0Stride, const uint16_t *pCamId1,
                                  uint32_t camId1Stride, const uint16_t *pCamId2, uint32_t camId2Stride)
{
    LoomArgs a;
    int rc = loom_args(a, width, height, count, pWeights, weightStride, pCamId0, camId0Stride, pCamId1, camId1Stride, pCamId2,
                       camId2Stride);
    if (rc) return rc;
    if (!pDst || !pCameras || !cameraStrides || !pMaps || dstStride < 3 * width || mapStride < 4 * width || (mapStride & 3u) ||
        ((uintptr_t)pMaps & 3u) || srcWidth < 2 || srcHeight < 2 || srcWidth >= 8192 || srcHeight >= 8192)
        return MVX_ERROR_INVALID_PARAMETERS;
    RemapSrc s{};
    for (uint32_t c = 0; c < count; c++) {
        if (!pCameras[c] || cameraStrides[c] < 3 * srcWidth) return MVX_ERROR_INVALID_PARAMETERS;
        s.cam[c] = pCameras[c];
        s.camStride[c] = cameraStrides[c];
    }
    s.map = (const uint8_t *)pMaps;
    s.mapStride = mapStride;
    s.height = height;
    s.sw = srcWidth;
    s.sh = srcHeight;
    const dim3 block(64, 4), grid((width + 63) / 64, (height + 3) / 4);
    hipLaunchKernelGGL(k_stitch_loom_fused, grid, block, 0, stream_of(stream), pDst, dstStride, a, s);
    MVX_LAUNCH_CHECK();
    return MVX_SUCCESS;
}
