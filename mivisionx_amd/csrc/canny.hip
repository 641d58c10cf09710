// canny.hip -- CannyEdgeDetector (gradient 3, 5, 7; L1 and L2 norms) for gfx950: Sobel -> non-max
// suppression with double threshold -> hysteresis.
//
// The reference traces hysteresis with a sequential LIFO stack on the CPU (its HIP backend has no
// trace kernel).  The traced image only depends on connectivity: a weak pixel ends at 255 iff it is
// 8-connected through weak pixels to a strong one.  Here that set is computed with a lock-free
// union-find over the edge pixels (atomicMin linking, then one find per pixel), three passes with no
// host round trip, in the 4-byte-per-pixel workspace the reference reserves for its XY stack.
#include "arith.hpp"
#include "rowmask.hpp"

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

// ---- gradient 5 / 7 and L2 Sobel planes (ago_haf_cpu_canny.cpp:145-427, 646-1056) -------------------
// 16-bit lane arithmetic of the reference's SSE code
MVX_DEV int w16(int v) { return (int16_t)(uint16_t)(uint32_t)v; }
MVX_DEV int s16(int v) { return min(max(v, -32768), 32767); }
MVX_DEV int sll16(int v, int k) { return w16((int)((uint32_t)(uint16_t)v << k)); }
MVX_DEV int abs16(int v) { return w16(v < 0 ? -v : v); }
MVX_DEV uint16_t canny_out(int mag16, uint32_t sector) { return (uint16_t)(sll16(mag16, 2) | (int)(sector & 3u)); }

// floor(sqrt(F)) for F < 2^32, exact: a float estimate corrected in integer arithmetic
MVX_DEV uint32_t isqrt_u64(uint64_t F)
{
    uint32_t k = (uint32_t)sqrtf((float)F);
    while ((uint64_t)k * k > F) k--;
    while ((uint64_t)(k + 1) * (k + 1) <= F) k++;
    return k;
}

// (vx_int16)sqrt((Gx*Gx) + (Gy*Gy)): 32-bit int sum, correctly rounded double sqrt truncated -- the
// integer square root -- with 0x80000000 for a negative (wrapped) sum, then 16 bits
MVX_DEV int l2_scalar(int gx, int gy)
{
    const int32_t q = (int32_t)((uint32_t)(gx * gx) + (uint32_t)(gy * gy));
    const int32_t t = q < 0 ? INT32_MIN : (int32_t)isqrt_u64((uint64_t)q);
    return (int16_t)(uint16_t)(uint32_t)t;
}

template <int G> struct Taps;
template <> struct Taps<3> { static constexpr int sm[3] = {1, 2, 1}, dv[3] = {-1, 0, 1}; };
template <> struct Taps<5> { static constexpr int sm[5] = {1, 4, 6, 4, 1}, dv[5] = {-1, -2, 0, 2, 1}; };
template <> struct Taps<7> { static constexpr int sm[7] = {1, 6, 15, 20, 15, 6, 1}, dv[7] = {-1, -4, -5, 0, 5, 4, 1}; };

// the scalar loop's value at (x, y), linear-address reads
template <int G, bool L2> MVX_DEV uint16_t canny_scalar(const uint8_t *src, long n, long stride, int x, int y)
{
    constexpr int R = G / 2;
    int gxs = 0, gys = 0;
#pragma unroll
    for (int j = 0; j < G; j++) {
        const long b = (long)(y - R + j) * stride + x;
#pragma unroll
        for (int i = 0; i < G; i++) {
            const int p = (int)ld_u8(src, n, b + i - R);
            gxs += Taps<G>::sm[j] * Taps<G>::dv[i] * p;
            // 5x5 L2: the first row's term lacks its parentheses (ago_haf_cpu_canny.cpp:880)
            const int wy = (G == 5 && L2 && j == 0 && i > 0) ? Taps<G>::sm[i] : Taps<G>::dv[j] * Taps<G>::sm[i];
            gys += wy * p;
        }
    }
    const int gx = w16(gxs);
    int gy = w16(gys);
    if (G == 5 && !L2) gy = w16(-gy);
    const int mag = L2 ? l2_scalar(gx, gy) : w16(abs(gx) + abs(gy));
    return canny_out(mag, canny_sector(gx, gy));
}

// The SSE passes over [0, aligned), aligned = w & ~15 >= 16: scratch rows r0, r1 (r1 = r0 + ALIGN16(w))
// with zeros before r0, after r1 and in their unwritten tails; a read past the end of r0 lands in r1
// (and one before r1 in r0) when w is a multiple of 16.  Each scratch slot x0-R .. stands for one source
// column (or none) for every row.  One wave owns a strip of 64 - 2R output columns and CSW_ROWS rows:
// lane l holds the slot of column x0 - R + l and keeps that column's G source bytes in a rolling
// registers (all of the strip's rows are loaded first); its two vertical sums reach the neighbouring lanes by DPP
// lane shifts, so no LDS and no barrier is involved.  Lanes R .. 63-R finish one column each; columns
// aligned-1 .. w-1 take the scalar loop's value.
// Short waves: a wave's rows are one chain of loads and dependent arithmetic, and many short waves keep
// more of the image in flight than a few long ones (3840x2160, gradient 5, L2: 2 rows per wave 28 us,
// 4 rows 29 us, 8 rows 35 us, 16 rows 50 us; 8 strips per workgroup 28 us, 4 strips 36 us).
constexpr int CSW_ROWS = 2;    // output rows per wave
constexpr int CSW_WAVES = 8;   // strips (waves) per workgroup

// which scratch value a slot holds: kind 1 = (r0, r1) of source column sc, 2 = r0 <- r1 of sc (the slot
// lies past the end of r0), 3 = r1 <- r0 of sc (before the start of r1), 0 = zeros
MVX_DEV int canny_slot(int col, int aligned, int a16, int &sc)
{
    if (col >= 0 && col < aligned) { sc = col; return 1; }
    if (col >= a16 && col - a16 < aligned) { sc = col - a16; return 2; }
    if (col < 0 && a16 + col < aligned) { sc = a16 + col; return 3; }
    sc = 0;
    return 0;
}

// the vertical pass of the SSE loop over one column's G source bytes: t0 = the smoothing sum (scratch
// row r0), t1 = the derivative sum (r1); gradient 7 shifts both right by 2.  |t0| <= 4080 and
// |t1| <= 765 for every G, so the reference's 16-bit scratch values are these.
template <int G> MVX_DEV void canny_vsum(const int (&s)[G], int &t0, int &t1)
{
    if (G == 3) {
        t0 = s[0] + s[2] + 2 * s[1];
        t1 = s[2] - s[0];
    } else if (G == 5) {
        t0 = 4 * (s[1] + s[3]) + 6 * s[2] + (s[0] + s[4]);
        t1 = 2 * (s[3] - s[1]) + (s[4] - s[0]);
    } else {
        t0 = (6 * (s[1] + s[5]) + 20 * s[3] + (s[0] + s[6]) + 15 * (s[2] + s[4])) >> 2;
        t1 = (4 * (s[5] - s[1]) + (s[6] - s[0]) + 5 * (s[4] - s[2])) >> 2;
    }
}

// the value of lane + 1 / lane - 1 (DPP wave_shl:1 / wave_shr:1; lanes past the ends read 0)
MVX_DEV int dpp_next(int v) { return __builtin_amdgcn_update_dpp(0, v, 0x130, 0xf, 0xf, true); }
MVX_DEV int dpp_prev(int v) { return __builtin_amdgcn_update_dpp(0, v, 0x138, 0xf, 0xf, true); }

// the SSE horizontal pass and packing of one column from the scratch slots at offsets -R..R
// (A0(d) = r0 slot, A1(d) = r1 slot)
template <int G, bool L2> MVX_DEV uint16_t canny_simd_px(const int (&a0)[G], const int (&a1)[G])
{
    constexpr int R = G / 2;
#define A0(d) (a0[R + (d)])
#define A1(d) (a1[R + (d)])
    int gx, gy;
    if (G == 3) {
        gx = w16(A0(1) - A0(-1));
        gy = w16(-w16(w16(A1(-1) + A1(1)) + sll16(A1(0), 1)));
    } else if (G == 5) {
        gx = s16(sll16(w16(A0(1) - A0(-1)), 1) + w16(A0(2) - A0(-2)));
        int t = w16(sll16(w16(A1(-1) + A1(1)), 2) + w16(A1(-2) + A1(2)));
        t = L2 ? s16(t + w16(A1(0) * 6)) : w16(t + w16(A1(0) * 6));
        gy = w16(-t);
    } else {
        int t = sll16(s16(A0(2) - A0(-2)), 2);
        const int u = w16(s16(A0(1) - A0(-1)) * 5);
        t = s16(t + s16(A0(3) - A0(-3)));
        gx = s16(t + u);
        int v = s16(w16(w16(A1(-2) + A1(2)) * 6) + w16(A1(0) * 20));
        const int q = w16(w16(A1(-1) + A1(1)) * 15);
        v = s16(v + s16(A1(-3) + A1(3)));
        v = s16(v + q);
        gy = s16(-v);
    }
#undef A0
#undef A1
    int mag;
    if (!L2) {
        mag = w16(abs16(gx) + abs16(gy));
        if (G == 7 && mag > 0x3fff) mag = 0x3fff;
    } else if (G == 3) {
        mag = l2_pack_simd((float)(uint16_t)w16(gx * gx) + (float)(uint16_t)w16(gy * gy));
    } else {
        mag = l2_pack_simd((float)(gx * gx + gy * gy));
    }
    return canny_out(mag, canny_sector(gx, gy));
}

// one strip: EDGE strips hold slots outside [0, aligned) or columns from aligned - 1 on
template <int G, bool L2, bool EDGE>
MVX_DEV void canny_strip(uint32_t w, uint32_t h, uint16_t *dst, uint32_t dstStride, const uint8_t *src, uint32_t srcStride,
                         int xs, int yb, int ye)
{
    constexpr int R = G / 2;
    const int lane = (int)threadIdx.x & 63;
    const int aligned = (int)(w & ~15u), a16 = (int)((w + 15u) & ~15u);
    const int col = xs - R + lane;
    int sc = col;
    const int kind = EDGE ? canny_slot(col, aligned, a16, sc) : 1;
    const __amdgpu_buffer_rsrc_t sR = plane_rsrc(src), dR = plane_rsrc(dst);
    // a zero slot still loads an in-plane byte (column 0) and drops it, so the loads stay unconditional
    auto load = [&](int r) -> int {
        const int v = (int)__builtin_amdgcn_raw_buffer_load_b8(sR, sc, (int)((uint32_t)r * srcStride), 0);
        return (EDGE && kind == 0) ? 0 : v;
    };
    // every source row of the strip is loaded before the first is used: the loads are all in flight
    // together, and row i waits only for its own G rows
    constexpr int NR = CSW_ROWS + G - 1;
    int rw[NR];
#pragma unroll
    for (int j = 0; j < NR; j++) rw[j] = (yb - R + j < ye + R) ? load(yb - R + j) : 0;
    const int x = col;   // output column of lanes R .. 63 - R
    const bool outLane = lane >= R && lane < 64 - R && x < (int)w;
#pragma unroll
    for (int i = 0; i < CSW_ROWS; i++) {
        const int y = yb + i;
        if (y >= ye) break;
        int win[G];
#pragma unroll
        for (int j = 0; j < G; j++) win[j] = rw[i + j];
        int t0, t1;
        canny_vsum<G>(win, t0, t1);
        int a0[G], a1[G];
        a0[R] = t0;
        a1[R] = t1;
        if (EDGE) {
            a0[R] = kind == 1 ? t0 : kind == 2 ? t1 : 0;
            a1[R] = kind == 1 ? t1 : kind == 3 ? t0 : 0;
        }
        // the lane shifts run with every lane active
#pragma unroll
        for (int d = 1; d <= R; d++) {
            a0[R + d] = dpp_next(a0[R + d - 1]);
            a0[R - d] = dpp_prev(a0[R - d + 1]);
            a1[R + d] = dpp_next(a1[R + d - 1]);
            a1[R - d] = dpp_prev(a1[R - d + 1]);
        }
        uint16_t out = canny_simd_px<G, L2>(a0, a1);
        if (outLane) {
            if (EDGE && x >= aligned - 1) out = canny_scalar<G, L2>(src, (long)h * srcStride, srcStride, x, y);
            __builtin_amdgcn_raw_buffer_store_b16(out, dR, 2 * x, (int)((uint32_t)y * dstStride), 0);
        }
    }
}

template <int G, bool L2>
__global__ void __launch_bounds__(64 * CSW_WAVES)
k_canny_sobel_g(uint32_t w, uint32_t h, uint16_t *__restrict__ dst, uint32_t dstStride, const uint8_t *__restrict__ src,
                uint32_t srcStride, uint32_t strips)
{
    constexpr int R = G / 2, CW = 64 - 2 * R;
    const int strip = (int)blockIdx.x * CSW_WAVES + ((int)threadIdx.x >> 6);
    if (strip >= (int)strips) return;   // the whole wave
    const int yb = R + (int)blockIdx.y * CSW_ROWS, ye = min(yb + CSW_ROWS, (int)h - R);
    const int xs = strip * CW, aligned = (int)(w & ~15u);
    if (xs - R < 0 || xs - R + 63 >= aligned)
        canny_strip<G, L2, true>(w, h, dst, dstStride, src, srcStride, xs, yb, ye);
    else
        canny_strip<G, L2, false>(w, h, dst, dstStride, src, srcStride, xs, yb, ye);
}

// widths below 16: only the scalar loop runs, from x = -1 (the slot before each row), rows in order
template <int G, bool L2>
__global__ void k_canny_sobel_g_narrow(uint32_t w, uint32_t h, uint16_t *dst, uint32_t dstStride, const uint8_t *src,
                                       uint32_t srcStride)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    constexpr int R = G / 2;
    const long n = (long)h * srcStride;
    for (int y = R; y < (int)h - R; y++)
        for (int x = -1; x < (int)w; x++)
            *(uint16_t *)((uint8_t *)dst + (long)y * dstStride + 2L * x) = canny_scalar<G, L2>(src, n, srcStride, x, y);
}

// CannySuppThreshold_U8XY_U16_3x3 (ago_haf_cpu_canny.cpp:549-596): x in [1, w-2], y in [1, h-2].
// A workgroup owns 256 columns x 16 output rows: the magnitude rows around them are staged in LDS
// with 16-byte loads (one dword per value, so no LDS load is ever misaligned), then each thread walks
// its column down the rows with a 3x3 window in registers.
constexpr int SUP_W = 256, SUP_H = 16, SUP_P = SUP_W + 16;   // LDS row: columns x0-8 .. x0+263

__global__ void __launch_bounds__(256)
k_canny_supp(uint32_t w, uint32_t h, uint8_t *__restrict__ dst, uint32_t dstStride, const uint16_t *__restrict__ mag,
             uint32_t magStride, uint32_t lower, uint32_t upper)
{
    __shared__ uint32_t t[(SUP_H + 2) * SUP_P];
    const int tid = (int)threadIdx.x;
    const int x0 = (int)blockIdx.x * SUP_W, y0 = 1 + (int)blockIdx.y * SUP_H;   // output rows y0 .. y0+15
    constexpr int CH = SUP_P / 8;                                               // 16-byte chunks per row
    for (int i = tid; i < (SUP_H + 2) * CH; i += 256) {
        const int r = i / CH, c = i - r * CH;
        const int y = y0 - 1 + r, col = x0 - 8 + 8 * c;
        uint4 v = make_uint4(0u, 0u, 0u, 0u);
        // a chunk starting inside the row lies inside its ALIGN16 bytes, which the stride covers
        if (y < (int)h && col >= 0 && col < (int)w)
            v = *reinterpret_cast<const uint4 *>((const uint8_t *)mag + (long)y * magStride + 2L * col);
        uint32_t *d = t + r * SUP_P + 8 * c;
        d[0] = v.x & 0xffffu; d[1] = v.x >> 16; d[2] = v.y & 0xffffu; d[3] = v.y >> 16;
        d[4] = v.z & 0xffffu; d[5] = v.z >> 16; d[6] = v.w & 0xffffu; d[7] = v.w >> 16;
    }
    __syncthreads();
    const int x = x0 + tid, cx = tid + 8;
    if (x < 1 || x > (int)w - 2) return;
    const int rows = min(SUP_H, (int)h - 1 - y0);
    // window: a = row y-1, b = row y, c = row y+1; index 0, 1, 2 = columns x-1, x, x+1
    uint32_t a0 = t[cx - 1], a1 = t[cx], a2 = t[cx + 1];
    uint32_t b0 = t[SUP_P + cx - 1], b1 = t[SUP_P + cx], b2 = t[SUP_P + cx + 1];
    uint8_t *out = dst + (long)y0 * dstStride + x;
    for (int r = 1; r <= rows; r++) {
        const uint32_t *n = t + (r + 1) * SUP_P + cx;
        const uint32_t c0 = n[-1], c1 = n[0], c2 = n[1];
        const int m = (int)(b1 >> 2), s = (int)(b1 & 3u);
        // neighbour pairs by sector (ago_haf_cpu_canny.cpp:26-31): 0 (-1,0)/(1,0), 1 (1,-1)/(-1,1),
        // 2 (0,-1)/(0,1), 3 (-1,-1)/(1,1)
        const uint32_t n0 = s == 0 ? b0 : s == 1 ? a2 : s == 2 ? a1 : a0;
        const uint32_t n1 = s == 0 ? b2 : s == 1 ? c0 : s == 2 ? c1 : c2;
        const int edge = (m > (int)(n0 >> 2) && m > (int)(n1 >> 2)) ? m : 0;
        *out = edge > (int)upper ? 255 : (edge <= (int)lower ? 0 : 127);
        out += dstStride;
        a0 = b0; a1 = b1; a2 = b2;
        b0 = c0; b1 = c1; b2 = c2;
    }
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

// ---------------------------------------------------------------- tiled edge trace ----------------
// HafCpu_CannyEdgeTrace_U8_U8XY (ago_haf_cpu_canny.cpp:598-645): every 127 pixel 8-connected to a 255
// pixel through 127 pixels becomes 255, every other 127 becomes 0.  Per 64x16 tile (one wave) the
// edge (127 or 255) and strong (255) pixels of the interior become 64-bit row masks (lane = row); two
// dilations run side by side in the wave, one from the strong pixels and one from the tile-border edge
// pixels.  Pixels the strong dilation reaches are final 255; edge pixels reached by neither are final
// 0; the rest (components that touch the tile border and hold no strong pixel) stay 127 and get a
// union-find id, the k-th such component of tile t the id 1 + t*CT_NB + k (at most one per border
// pixel).  Border tables record per border pixel 0 (the strong node), an id, or -1.  A second pass
// unions the tables across tile edges, a third re-finds the undecided components (the 127 pixels left
// in the tile, discovered in the same order) and writes 255 where the root is the strong node, else 0.
constexpr int CT_W = 64, CT_H = 16, CT_NB = 2 * (CT_W + CT_H);

struct CtArgs {
    uint32_t w, h, stride, tilesX, tilesY, tiles;
    uint8_t *img;
    uint32_t *parent;   // 1 + tiles * CT_NB: slot 0 the strong node (its own root), then CT_NB per tile
    int32_t *border;    // tiles * CT_NB: top row, bottom row (CT_W each), left, right column (CT_H each)
    uint32_t *ucount;   // tiles: undecided components of the tile
};

struct CtTile { int tile, tx0, ty0, tw, th; };

MVX_DEV bool ct_tile(const CtArgs &a, CtTile &t)
{
    t.tile = (int)blockIdx.x * 4 + __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    if (t.tile >= (int)a.tiles) return false;
    t.tx0 = (t.tile % (int)a.tilesX) * CT_W;
    t.ty0 = (t.tile / (int)a.tilesX) * CT_H;
    t.tw = min(CT_W, (int)a.w - t.tx0);
    t.th = min(CT_H, (int)a.h - t.ty0);
    return true;
}

// the tile's interior pixels, lane = column, one register per row (0 outside the interior)
MVX_DEV void ct_load(const CtArgs &a, const CtTile &t, int lane, uint32_t (&v)[CT_H])
{
    const int x = t.tx0 + lane;
    const bool colOk = lane < t.tw && x >= 1 && x <= (int)a.w - 2;
    gptr_t<const uint8_t> img = gmem((const uint8_t *)a.img);
#pragma unroll
    for (int r = 0; r < CT_H; r++) {
        const int y = t.ty0 + r;
        v[r] = (colOk && r < t.th && y >= 1 && y <= (int)a.h - 2) ? (uint32_t)img[(long)y * a.stride + x] : 0u;
    }
}

// lane r < CT_H: the mask of row r of the pixels equal to `value` (or, with value2, to either)
MVX_DEV uint64_t ct_rows(const uint32_t (&v)[CT_H], int lane, uint32_t value, uint32_t value2)
{
    uint64_t m = 0;
#pragma unroll
    for (int r = 0; r < CT_H; r++) {
        const uint64_t b = __ballot(v[r] == value || v[r] == value2);
        if (lane == r) m = b;
    }
    return m;
}

__global__ void __launch_bounds__(256) k_ct_tile(CtArgs a)
{
    const int lane = (int)threadIdx.x & 63;
    CtTile t;
    if (!ct_tile(a, t)) return;
    uint32_t v[CT_H];
    ct_load(a, t, lane, v);
    const uint64_t em = ct_rows(v, lane, 127u, 255u), sm = ct_rows(v, lane, 255u, 255u);
    const int rr = lane & 31;
    const uint64_t e = __shfl(em, rr);
    uint64_t seeds;
    if (lane < 32) seeds = __shfl(sm, rr);
    else seeds = (rr == 0 || rr == t.th - 1) ? e : (e & (1ull | (1ull << (t.tw - 1))));
    const uint64_t grown = grow_rows(seeds, e, lane);
    const uint64_t other = __shfl(grown, lane ^ 32);
    const uint64_t rm = lane < 32 ? grown : other;                  // lane r: row r reached from a strong pixel
    const uint64_t bmAll = lane < 32 ? (other & ~grown) : 0ull;     // lane r < 32: row r undecided
    // settle the 127 pixels: reached -> 255, in no border-touching component -> 0, undecided stay
    gptr_t<uint8_t> img = gmem(a.img);
#pragma unroll
    for (int r = 0; r < CT_H; r++) {
        if (r < t.th) {
            const uint64_t R = readlane64(rm, r), B = readlane64(bmAll, r);
            const bool reached = (R >> lane) & 1ull, undecided = (B >> lane) & 1ull;
            if (v[r] == 127u && (reached || !undecided))
                img[(long)(t.ty0 + r) * a.stride + t.tx0 + lane] = reached ? 255 : 0;
        }
    }
    int32_t topv = -1, botv = -1, leftv = -1, rightv = -1;
    {
        const uint64_t R0 = readlane64(rm, 0), Rb = readlane64(rm, t.th - 1);
        if (lane < t.tw && ((R0 >> lane) & 1ull)) topv = 0;
        if (lane < t.tw && ((Rb >> lane) & 1ull)) botv = 0;
        if (lane < t.th && (rm & 1ull)) leftv = 0;
        if (lane < t.th && ((rm >> (t.tw - 1)) & 1ull)) rightv = 0;
    }
    gptr_t<uint32_t> par = gmem(a.parent);
    if (t.tile == 0 && lane == 0) par[0] = 0;
    uint32_t k = 0;
    uint64_t bm = bmAll;
    while (true) {
        const uint64_t rows = __ballot(bm != 0ull);
        if (!rows) break;
        const int r0 = __builtin_ctzll(rows);
        const int c0 = __builtin_ctzll(readlane64(bm, r0));
        const uint32_t gid = 1u + (uint32_t)t.tile * CT_NB + k;
        const uint64_t c = grow_rows(lane == r0 ? (1ull << c0) : 0ull, bm, lane);
        if (lane == 0) par[gid] = gid;
        const uint64_t C0 = readlane64(c, 0), Cb = readlane64(c, t.th - 1);
        if (lane < t.tw && ((C0 >> lane) & 1ull)) topv = (int32_t)gid;
        if (lane < t.tw && ((Cb >> lane) & 1ull)) botv = (int32_t)gid;
        if (lane < t.th && (c & 1ull)) leftv = (int32_t)gid;
        if (lane < t.th && ((c >> (t.tw - 1)) & 1ull)) rightv = (int32_t)gid;
        bm &= ~c;
        k++;
    }
    gptr_t<int32_t> bt = gmem(a.border) + (long)t.tile * CT_NB;
    bt[lane] = topv;
    bt[CT_W + lane] = botv;
    if (lane < CT_H) { bt[2 * CT_W + lane] = leftv; bt[2 * CT_W + CT_H + lane] = rightv; }
    if (lane == 0) gmem(a.ucount)[t.tile] = k;
}

// the unions across the tile's right and bottom edges and its two lower corners
__global__ void __launch_bounds__(256) k_ct_border(CtArgs a)
{
    const int lane = (int)threadIdx.x & 63;
    CtTile t;
    if (!ct_tile(a, t)) return;
    const int tX = t.tile % (int)a.tilesX, tY = t.tile / (int)a.tilesX;
    gptr_t<const int32_t> tables = gmem((const int32_t *)a.border);
    gptr_t<const int32_t> bt = tables + (long)t.tile * CT_NB;
    gptr_t<uint32_t> par = gmem(a.parent);
    if (tX + 1 < (int)a.tilesX) {           // a right neighbour: this tile is full width
        gptr_t<const int32_t> nb = tables + (long)(t.tile + 1) * CT_NB;
        border_unions(par, lane < t.th ? bt[2 * CT_W + CT_H + lane] : -1, lane < t.th ? nb[2 * CT_W + lane] : -1, lane);
    }
    if (tY + 1 < (int)a.tilesY) {           // a lower neighbour: this tile is full height, same width
        gptr_t<const int32_t> nb = tables + (long)(t.tile + a.tilesX) * CT_NB;
        border_unions(par, lane < t.tw ? bt[CT_W + lane] : -1, lane < t.tw ? nb[lane] : -1, lane);
        if (lane == 0) {
            if (tX + 1 < (int)a.tilesX) {
                const int32_t p = bt[CT_W + t.tw - 1], q = tables[(long)(t.tile + a.tilesX + 1) * CT_NB];
                if (p >= 0 && q >= 0 && p != q) gunion(par, (uint32_t)p, (uint32_t)q);
            }
            if (tX > 0) {
                const int32_t p = bt[CT_W], q = tables[(long)(t.tile + a.tilesX - 1) * CT_NB + CT_W - 1];
                if (p >= 0 && q >= 0 && p != q) gunion(par, (uint32_t)p, (uint32_t)q);
            }
        }
    }
}

// the undecided components of each tile, found again in the tile pass's order, take their root's state
__global__ void __launch_bounds__(256) k_ct_resolve(CtArgs a)
{
    const int lane = (int)threadIdx.x & 63;
    CtTile t;
    if (!ct_tile(a, t)) return;
    if (gmem((const uint32_t *)a.ucount)[t.tile] == 0) return;
    uint32_t v[CT_H];
    ct_load(a, t, lane, v);
    uint64_t bm = ct_rows(v, lane, 127u, 127u);
    if (lane >= 32) bm = 0;
    gptr_t<uint8_t> img = gmem(a.img);
    uint32_t k = 0;
    while (true) {
        const uint64_t rows = __ballot(bm != 0ull);
        if (!rows) break;
        const int r0 = __builtin_ctzll(rows);
        const int c0 = __builtin_ctzll(readlane64(bm, r0));
        const uint64_t c = grow_rows(lane == r0 ? (1ull << c0) : 0ull, bm, lane);
        const uint32_t g = gfind(gmem((const uint32_t *)a.parent), 1u + (uint32_t)t.tile * CT_NB + k);
        const uint8_t val = g == 0 ? 255 : 0;
        uint64_t crow = __ballot(c != 0ull);
        while (crow) {
            const int r = __builtin_ctzll(crow);
            crow &= crow - 1;
            if ((readlane64(c, r) >> lane) & 1ull) img[(long)(t.ty0 + r) * a.stride + t.tx0 + lane] = val;
        }
        bm &= ~c;
        k++;
    }
}

struct CtLayout { size_t parent, border, ucount, total; uint32_t tilesX, tilesY, tiles; };

inline CtLayout ct_layout(uint32_t w, uint32_t h)
{
    CtLayout L;
    L.tilesX = (w + CT_W - 1) / CT_W;
    L.tilesY = (h + CT_H - 1) / CT_H;
    L.tiles = L.tilesX * L.tilesY;
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    L.parent = 0;
    L.border = al(((size_t)L.tiles * CT_NB + 1) * 4);
    L.ucount = L.border + al((size_t)L.tiles * CT_NB * 4);
    L.total = L.ucount + al((size_t)L.tiles * 4);
    return L;
}

}  // namespace mvx

using namespace mvx;

namespace {
template <int G, bool L2>
int canny_sobel_g(mvx_stream_t stream, uint32_t w, uint32_t h, uint16_t *dst, uint32_t dstStride, const uint8_t *src,
                  uint32_t srcStride)
{
    if (w == 0 || h == 0) return MVX_SUCCESS;
    if (!plane_ok_c(dst, w, dstStride, 2) || !plane_ok_c(src, w, srcStride, 1)) return MVX_ERROR_INVALID_PARAMETERS;
    if ((int)h < G) return MVX_SUCCESS;
    // buffer offsets (row * stride + column) stay below 2^32
    if ((uint64_t)h * dstStride > 0xffffffffull || (uint64_t)h * srcStride > 0xffffffffull) return MVX_ERROR_INVALID_PARAMETERS;
    if (w < 16) {
        hipLaunchKernelGGL((k_canny_sobel_g_narrow<G, L2>), dim3(1), dim3(64), 0, stream_of(stream), w, h, dst, dstStride, src, srcStride);
    } else {
        const uint32_t strips = (w + (64 - 2 * (G / 2)) - 1) / (64 - 2 * (G / 2));
        hipLaunchKernelGGL((k_canny_sobel_g<G, L2>), dim3((strips + CSW_WAVES - 1) / CSW_WAVES, (h - 2 * (G / 2) + CSW_ROWS - 1) / CSW_ROWS),
                           dim3(64 * CSW_WAVES), 0, stream_of(stream), w, h, dst, dstStride, src, srcStride, strips);
    }
    MVX_LAUNCH_CHECK();
    return MVX_SUCCESS;
}
}  // namespace

#define MVX_CANNY_SOBEL(NAME, G, L2)                                                                                      \
    extern "C" int NAME(mvx_stream_t stream, uint32_t w, uint32_t h, uint16_t *dst, uint32_t dstStride, const uint8_t *src, \
                        uint32_t srcStride)                                                                               \
    {                                                                                                                     \
        return canny_sobel_g<G, L2>(stream, w, h, dst, dstStride, src, srcStride);                                         \
    }
MVX_CANNY_SOBEL(mvx_CannySobel_U16_U8_5x5_L1NORM, 5, false)
MVX_CANNY_SOBEL(mvx_CannySobel_U16_U8_7x7_L1NORM, 7, false)
MVX_CANNY_SOBEL(mvx_CannySobel_U16_U8_3x3_L2NORM, 3, true)
MVX_CANNY_SOBEL(mvx_CannySobel_U16_U8_5x5_L2NORM, 5, true)
MVX_CANNY_SOBEL(mvx_CannySobel_U16_U8_7x7_L2NORM, 7, true)
#undef MVX_CANNY_SOBEL

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
    dim3 grid((w + SUP_W - 1) / SUP_W, (h - 2 + SUP_H - 1) / SUP_H);
    hipLaunchKernelGGL(k_canny_supp, grid, dim3(256), 0, stream_of(stream), w, h, dst, dstStride, mag, magStride,
                       lower, upper);
    MVX_LAUNCH_CHECK();
    return MVX_SUCCESS;
}

extern "C" int mvx_CannyEdgeTrace_U8_U8XY(mvx_stream_t stream, uint32_t w, uint32_t h, uint8_t *img, uint32_t stride,
                                          void *xyStack, uint32_t capacityOfXY)
{
    if (w == 0 || h == 0) return MVX_SUCCESS;
    // the stack holds 4-byte items (two 16-bit coordinates); both traces keep 32-bit words in it
    if (!plane_ok_c(img, w, stride, 1) || !xyStack || ((uintptr_t)xyStack & 3u) || (uint64_t)capacityOfXY < (uint64_t)w * h)
        return MVX_ERROR_INVALID_PARAMETERS;
    if (w < 3 || h < 3) return MVX_SUCCESS;
    const CtLayout L = ct_layout(w, h);
    if (L.total <= (size_t)capacityOfXY * 4) {
        // the tiled trace, its tables in the stack's memory (about 1.25 bytes per pixel)
        uint8_t *ws = (uint8_t *)xyStack;
        CtArgs a{w, h, stride, L.tilesX, L.tilesY, L.tiles, img, (uint32_t *)(ws + L.parent), (int32_t *)(ws + L.border),
                 (uint32_t *)(ws + L.ucount)};
        const dim3 grid((L.tiles + 3) / 4);
        hipLaunchKernelGGL(k_ct_tile, grid, dim3(256), 0, stream_of(stream), a);
        hipLaunchKernelGGL(k_ct_border, grid, dim3(256), 0, stream_of(stream), a);
        hipLaunchKernelGGL(k_ct_resolve, grid, dim3(256), 0, stream_of(stream), a);
        MVX_LAUNCH_CHECK();
        return MVX_SUCCESS;
    }
    // small images: a union-find over every interior edge pixel, the labels in the stack (4 bytes per pixel)
    dim3 grid((w - 2 + 255) / 256, h - 2);
    uint32_t *lab = (uint32_t *)xyStack;
    hipLaunchKernelGGL(k_canny_init, grid, dim3(256), 0, stream_of(stream), w, h, img, stride, lab);
    hipLaunchKernelGGL(k_canny_merge, grid, dim3(256), 0, stream_of(stream), w, h, img, stride, lab);
    hipLaunchKernelGGL(k_canny_mark, grid, dim3(256), 0, stream_of(stream), w, h, img, stride, lab);
    hipLaunchKernelGGL(k_canny_final, grid, dim3(256), 0, stream_of(stream), w, h, img, stride, lab);
    MVX_LAUNCH_CHECK();
    return MVX_SUCCESS;
}
