// vision_graph.hip -- the RGB vision graph as a few fused, frame-batched launches for gfx950.
//
// Graph (the reference schedules ~15 kernels per frame for it, with host round trips for the Canny
// trace and the Harris pick):
//   RGB -> ColorConvert Y (virtual) -> { ChannelCopy -> pyramid level 0,
//                                         ScaleImage 2:1 bilinear,
//                                         ScaleGaussianHalf_5x5 -> pyramid levels 1..,
//                                         Canny 3x3 L1 (Sobel -> NMS/threshold -> hysteresis),
//                                         Harris 3x3/3x3 (HG3 -> score -> NMS -> sort and pick) }
//
// Launches per batch of frames (blockIdx.z = frame):
//   k_front     one 58x32 luma tile (+3 halo) per workgroup of four waves.  Every window row of the
//               tile (64 luma, 62 gradient, 60 score columns) fits one 64-lane wave, so lane = column
//               and a wave walks rows: RGB -> Y in LDS, then level 0, the 2:1 scaled image, level 1,
//               Canny gradients/NMS/threshold with tile-local hysteresis (row-mask dilation; union-find
//               only over the weak components that touch the tile border), Harris products / exact
//               integer window sums / score / NMS with candidate append.  Only graph outputs and small
//               per-tile tables reach HBM.  Tiles away from the image border take a branch-free
//               instantiation.
//   k_border    unions the hysteresis components that touch tile borders, across tiles.  Strong
//               components get ids below every weak one, so after atomicMin linking a set's root is
//               strong iff the set holds a strong component -- no marking pass.
//   k_resolve   the undecided weak pixels k_front listed per tile take their global root's state.
//   k_half      pyramid levels 2.. (one launch per level, all frames).
//   k_pick      Harris sort and pick, one workgroup per frame (pick.hpp).
//
// Every output is bit-exact with the reference CPU path; the equivalences used are stated where they
// are used.  Limits of the fused path (callers use the per-node kernels otherwise): width >= 16 and a
// multiple of 4 (the HG3 pitch quirk is then the identity), even height, scaled image exactly half
// size, every pyramid step one the reference accepts.
#include "mvx_common.hpp"
#include "pick.hpp"
#include "pyramid.hpp"
#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

namespace mvx {

constexpr int TW = 58, TH = 32;                  // luma tile
constexpr int HY = 3;                            // luma halo: Sobel 1 + Harris window 1 + NMS 1
constexpr int TP = 64;                           // LDS row pitch (one wave row)
constexpr int YH = TH + 2 * HY;                  // luma window rows (columns: TW + 6 = 64)
constexpr int GH = TH + 4, GW = TW + 4;          // gradients, halo 2
constexpr int MH = TH + 2, MW = TW + 2;          // Canny magnitude / Harris score, halo 1
constexpr int TS = TH * TP;                      // tile-local pixel index space, p = ly * TP + lx
constexpr int FT = 256;                          // threads of the tile kernels (4 waves)
constexpr int NW = FT / 64;

struct FrameDesc {
    const uint8_t *rgb; uint32_t rgbStride;
    uint8_t *lvl[8]; uint32_t lvlStride[8];      // pyramid levels, lvl[0] = luma
    uint8_t *scaled; uint32_t scaledStride;
    uint8_t *edges; uint32_t edgesStride;
    void *corners; uint32_t *numItems; uint64_t *count;
    // per-frame workspace
    int32_t *border;      // tiles x 2(TW+TH) component ids of tile-border pixels (-1: none)
    uint32_t *parent;     // 1 + tiles * TS union-find slots (sparsely used)
    uint32_t *ulist;      // tiles x TS: a tile's undecided pixels, p | tile-local root << 16
    uint32_t *ucount;     // tiles: undecided pixels listed
    void *cand;           // Harris XYS candidates
    uint32_t *candCount;  // candidates appended (reset by k_pick)
    void *pickWs;         // harris pick workspace
};

struct GraphParams {
    int w, h, yStride, tilesX, tilesY;
    int dw, dh;            // scaled size (w/2, h/2)
    int levels;
    int lvlW[8], lvlH[8];
    uint32_t lower, upper;
    float k, thr, norm, dist;
    float normRcp;           // RN(1 / norm)
    uint32_t capacity, candCap;
};

// luma of one RGB pixel, HafCpu_ColorConvert_Y_RGB (ago_haf_cpu_color_convert.cpp:2913-2990):
// ((R * 0.2126f) + G * 0.7152f) + B * 0.0722f in float, truncated, saturated
template <class Ptr> MVX_DEV uint32_t luma(Ptr p)
{
    float v = (float)p[0] * 0.2126f;
    v = v + (float)p[1] * 0.7152f;
    v = v + (float)p[2] * 0.0722f;
    const int iv = (int)v;
    return (uint32_t)min(max(iv, 0), 255);
}

// the virtual luma plane at linear position (r, c): stride yStride, padding and guard bytes zero
MVX_DEV uint32_t yflat(const FrameDesc &F, const GraphParams &P, int r, int c)
{
    if (c >= 0 && c < P.w && r >= 0 && r < P.h) return luma(gmem(F.rgb) + (long)r * F.rgbStride + 3L * c);
    const long i = (long)r * P.yStride + c;
    if (i < 0 || i >= (long)P.h * P.yStride) return 0;
    r = (int)(i / P.yStride);
    c = (int)(i - (long)r * P.yStride);
    if (c >= P.w) return 0;
    return luma(gmem(F.rgb) + (long)r * F.rgbStride + 3L * c);
}

// HafCpu_FastAtan2_Canny (ago_haf_cpu_arithmetic.cpp:9638-9656), statement order kept
MVX_DEV uint32_t canny_sector_f(int gx, int gy)
{
    const uint32_t ax = (uint32_t)abs(gx) & 0xffffu, ay = (uint32_t)abs(gy) & 0xffffu;
    const float d1 = (float)ax * 0.4142135623730950488016887242097f;
    const float d2 = (float)ax * 2.4142135623730950488016887242097f;
    uint32_t ret = (gx * gy) < 0 ? 3u : 1u;
    if ((float)ay <= d1) ret = 0;
    if ((float)ay >= d2) ret = 2;
    return ret;
}

MVX_DEV uint16_t canny_pack_f(int gx, int gyn)
{
    int16_t t = (int16_t)(abs((int)(int16_t)gx) + abs((int)(int16_t)gyn));
    t = (int16_t)(t << 2);
    t = (int16_t)(t | (int16_t)(canny_sector_f((int16_t)gx, (int16_t)gyn) & 3u));
    return (uint16_t)t;
}

// the value of lane + 1 of the wave (DPP wave_shl:1; lane 63 reads 0)
MVX_DEV int lane_next(int v) { return __builtin_amdgcn_update_dpp(0, v, 0x130, 0xf, 0xf, true); }

// union-find with atomicMin linking (the larger root hangs under the smaller): the final root of
// a component is its smallest index, whatever the order of the unions
MVX_DEV int lfind(const int *L, int p)
{
    int l = L[p];
    while (l != p) { p = l; l = L[p]; }
    return p;
}

MVX_DEV void lunion(int *L, int a, int b)
{
    while (true) {
        a = lfind(L, a);
        b = lfind(L, b);
        if (a == b) return;
        if (a < b) { int t = a; a = b; b = t; }
        const int old = atomicMin(&L[a], b);
        if (old == a) return;
        a = old;
    }
}

MVX_DEV uint32_t gfind(gptr_t<const uint32_t> par, uint32_t p)
{
    uint32_t l = par[p];
    while (l != p) { p = l; l = par[p]; }
    return p;
}

MVX_DEV void gunion(gptr_t<uint32_t> par, uint32_t a, uint32_t b)
{
    while (true) {
        a = gfind(par, a);
        b = gfind(par, b);
        if (a == b) return;
        if (a < b) { uint32_t t = a; a = b; b = t; }
        const uint32_t old = __hip_atomic_fetch_min(par + a, b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (old == a) return;
        a = old;
    }
}

// Arrays whose lifetimes do not overlap share LDS (every hand-over is separated by a barrier):
// the gradients (steps 2-3) with the row masks of the Canny states (steps 4-6), the Harris score
// (steps 3-4 and the candidate append) with the union-find labels (step 6).  Under 25 KB per
// workgroup: 6 resident workgroups per CU.
struct FrontShared {
    uint8_t Y[YH][TP];
    union {
        struct { int16_t Gx[GH][TP], Gy[GH][TP]; };
        // per row (bit = column): edge (127/255), strong (255), reached from a strong pixel, and
        // weak pixels of the components that touch the tile border but hold no strong pixel
        struct { uint64_t Em[TH], Sm[TH], Rm[TH], Bm[TH]; };
    };
    uint16_t M[MH][TP];
    union {
        float V[MH][TP];
        int L[TS];                             // union-find labels of the weak pixels no strong pixel reaches
    };
    uint32_t cnt, base, unres;   // candidates, their global offset, undecided pixels listed
};

// INNER: the whole luma window lies inside the image and every Canny / Harris position of the tile
// window is an interior pixel, so no image-border rule applies.
template <bool INNER>
MVX_DEV void front_tile(FrontShared &s, const FrameDesc &F, const GraphParams &P, int tx0, int ty0, int tile)
{
    const int t = (int)threadIdx.x, lane = t & 63, wave = t >> 6;
    const int w = P.w, h = P.h;
    const int tw = INNER ? TW : min(TW, w - tx0), th = INNER ? TH : min(TH, h - ty0);

    // 1. luma window: lane = column (tx0-3+lane), linear-address semantics of the virtual Y plane.
    //    Inner tiles issue all of a lane's RGB loads before converting, so their latencies overlap.
    if (INNER) {
        constexpr int RPW = (YH + NW - 1) / NW;   // rows per wave
        uint8_t px[RPW][3];
        gptr_t<const uint8_t> col = gmem(F.rgb) + (long)(ty0 - HY + wave) * F.rgbStride + 3L * (tx0 - HY + lane);
#pragma unroll
        for (int k = 0; k < RPW; k++) {
            if (wave + k * NW < YH) {
                gptr_t<const uint8_t> p = col + (long)(k * NW) * F.rgbStride;
                px[k][0] = p[0]; px[k][1] = p[1]; px[k][2] = p[2];
            }
        }
#pragma unroll
        for (int k = 0; k < RPW; k++)
            if (wave + k * NW < YH) s.Y[wave + k * NW][lane] = (uint8_t)luma(px[k]);
    } else {
        for (int r = wave; r < YH; r += NW) {
            const int y = ty0 - HY + r, x = tx0 - HY + lane;
            s.Y[r][lane] = (uint8_t)yflat(F, P, y, x);
        }
    }
    __syncthreads();

    // 2. Sobel gradients (linear-address reads) at (ty0-2+r, tx0-2+lane)
    if (lane < GW) {
        for (int r = wave; r < GH; r += NW) {
            const int a0 = s.Y[r][lane], a1 = s.Y[r][lane + 1], a2 = s.Y[r][lane + 2];
            const int b0 = s.Y[r + 1][lane], b2 = s.Y[r + 1][lane + 2];
            const int c0 = s.Y[r + 2][lane], c1 = s.Y[r + 2][lane + 1], c2 = s.Y[r + 2][lane + 2];
            s.Gx[r][lane] = (int16_t)((c2 - c0) + (a2 - a0) + ((b2 - b0) << 1));
            s.Gy[r][lane] = (int16_t)((c0 + (c1 << 1) + c2) - (a0 + (a1 << 1) + a2));
        }
    }
    __syncthreads();

    // 3. at (ty0-1+r, tx0-1+lane): Canny U16 magnitude|sector; Harris score from a per-column
    //    sliding window of 3-column sums of the HG3 products {gx^2, gx gy, gy^2}.  HG3 rows 0 and h-1
    //    are never written (zero).  The products are integers whose 9-sums stay below 2^24, so int32
    //    sums equal the reference's float sums exactly.
    //    Each lane forms the products of its own gradient column; the 3-column sums (columns
    //    lane..lane+2) take the right neighbours' through DPP wave shifts, so this part runs with all
    //    64 lanes active (lanes >= MW compute values nobody stores).
    {
        const int x = tx0 - 1 + lane;
        const int r0 = wave * ((MH + NW - 1) / NW), r1 = min(r0 + (MH + NW - 1) / NW, MH);
        int hxx[3], hxy[3], hyy[3];
        auto hsum = [&](int gr, int &sxx, int &sxy, int &syy) {
            int pxx = 0, pxy = 0, pyy = 0;
            if (INNER || (ty0 - 2 + gr >= 1 && ty0 - 2 + gr <= h - 2)) {
                const int gx = s.Gx[gr][lane], gy = s.Gy[gr][lane];
                pxx = gx * gx; pxy = gx * gy; pyy = gy * gy;
            }
            sxx = pxx + lane_next(pxx + lane_next(pxx));
            sxy = pxy + lane_next(pxy + lane_next(pxy));
            syy = pyy + lane_next(pyy + lane_next(pyy));
        };
        hsum(r0, hxx[0], hxy[0], hyy[0]);
        hsum(r0 + 1, hxx[1], hxy[1], hyy[1]);
        for (int r = r0; r < r1; r++) {
            const int y = ty0 - 1 + r;
            hsum(r + 2, hxx[2], hxy[2], hyy[2]);
            if (lane < MW) {
                // Canny
                uint16_t m = 0;
                if (INNER || (y >= 1 && y <= h - 2 && x >= 1 && x < w)) {
                    m = canny_pack_f(s.Gx[r + 1][lane + 1], -(int)s.Gy[r + 1][lane + 1]);
                } else if (y >= 1 && y <= h - 2 && x == 0) {
                    // CannySobel column 0 (ago_haf_cpu_canny.cpp:83-116 with the staged rows r0[-1] = 0 and
                    // r1[-1] = r0[ALIGN16(w)-1]); luma window column 3 is x = 0, rows r+1..r+3 are y-1..y+1
                    const int t0_1 = (int)s.Y[r + 1][4] + (int)s.Y[r + 3][4] + 2 * (int)s.Y[r + 2][4];
                    const int t1_0 = (int)s.Y[r + 3][3] - (int)s.Y[r + 1][3];
                    const int t1_1 = (int)s.Y[r + 3][4] - (int)s.Y[r + 1][4];
                    const int t1_m1 = (w % 16 == 0) ? ((int)yflat(F, P, y - 1, w - 1) + (int)yflat(F, P, y + 1, w - 1) +
                                                       2 * (int)yflat(F, P, y, w - 1)) : 0;
                    m = canny_pack_f(t0_1, -(t1_m1 + 2 * t1_0 + t1_1));
                }
                s.M[r][lane] = m;
                // Harris
                float v = 0.0f;
                if (INNER || (x >= 1 && x <= w - 2 && y >= 1 && y <= h - 2)) {
                    const float gx2 = (float)(hxx[0] + hxx[1] + hxx[2]);
                    const float gxy2 = (float)(hxy[0] + hxy[1] + hxy[2]);
                    const float gy2 = (float)(hyy[0] + hyy[1] + hyy[2]);
                    const float tr = gx2 + gy2;
                    const float det = (gx2 * gy2) - (gxy2 * gxy2);
                    const float mc = det - (P.k * tr * tr);
                    // Vc = mc / norm > T ? mc / norm : 0, the quotient exactly the IEEE one
                    const float q = div_by_const_rn(mc, P.norm, P.normRcp);
                    v = (q > P.thr) ? q : 0.0f;
                }
                s.V[r][lane] = v;
            }
            hxx[0] = hxx[1]; hxy[0] = hxy[1]; hyy[0] = hyy[1];
            hxx[1] = hxx[2]; hxy[1] = hxy[2]; hyy[1] = hyy[2];
        }
    }
    __syncthreads();

    // 4. tile pixels (ty0+r, tx0+lane): Canny NMS + double threshold, Harris NMS -> candidates
    uint32_t myCand = 0;                 // bit q: row wave + q*NW is a candidate
    for (int q = 0; q < TH / NW; q++) {
        const int r = wave + q * NW;
        uint8_t st = 0;
        bool cand = false;
        const int y = ty0 + r, x = tx0 + lane;
        if (lane < TW && (INNER || (y >= 1 && y <= h - 2 && x >= 1 && x <= w - 2))) {
            const int j = r + 1, i = lane + 1;
            const uint32_t mv = s.M[j][i];
            const int m = (int)(mv >> 2), a = (int)(mv & 3u);
            const int dx0 = (a == 0) ? -1 : (a == 1) ? 1 : (a == 2) ? 0 : -1;
            const int dy0 = (a == 0) ? 0 : -1;
            const int n0 = (int)(s.M[j + dy0][i + dx0] >> 2), n1 = (int)(s.M[j - dy0][i - dx0] >> 2);
            const int edge = (m > n0 && m > n1) ? m : 0;
            st = edge > (int)P.upper ? 255 : (edge <= (int)P.lower ? 0 : 127);
            const float c = s.V[j][i];
            cand = c >= s.V[j - 1][i - 1] && c >= s.V[j - 1][i] && c >= s.V[j - 1][i + 1] && c >= s.V[j][i - 1] &&
                   c > s.V[j][i + 1] && c > s.V[j + 1][i - 1] && c > s.V[j + 1][i] && c > s.V[j + 1][i + 1];
        }
        if (cand) myCand |= 1u << q;
        // row masks of edge (127/255) and strong (255) pixels: bit = column
        const uint64_t em = __ballot(st != 0), sm = __ballot(st == 255);
        if (lane == 0) { s.Em[r] = em; s.Sm[r] = sm; }
    }
    // candidate append: one LDS atomic per thread with candidates, one global atomic per workgroup
    const uint32_t nCand = (uint32_t)__popc(myCand);
    uint32_t myOff = 0;
    if (nCand) myOff = atomicAdd(&s.cnt, nCand);
    __syncthreads();
    if (t == 0 && s.cnt) s.base = __hip_atomic_fetch_add(gmem(F.candCount), s.cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    if (nCand) {
        // one XYS record {int16 x, int16 y, float s} as a single 64-bit store
        static_assert(sizeof(XYS) == 8, "XYS layout");
        gptr_t<uint64_t> list = gmem((uint64_t *)F.cand);
        uint32_t pos = s.base + myOff;
        for (int q = 0; q < TH / NW; q++) {
            if (!(myCand & (1u << q))) continue;
            const int r = wave + q * NW;
            if (pos < P.candCap) {
                list[pos] = (uint64_t)(uint16_t)(tx0 + lane) | ((uint64_t)(uint16_t)(ty0 + r) << 16) |
                            ((uint64_t)__float_as_uint(s.V[r + 1][lane + 1]) << 32);
            }
            pos++;
        }
    }

    // 5. outputs that only need luma: level 0, the 2:1 scaled image, level 1
    if (lane < tw) {
        gptr_t<uint8_t> l0 = gmem(F.lvl[0]) + (long)ty0 * F.lvlStride[0] + tx0 + lane;
        for (int r = wave; r < th; r += NW) l0[(long)r * F.lvlStride[0]] = s.Y[r + HY][lane + HY];
    }
    if (lane < TW / 2) {
        // ScaleImage 2:1: xm = 2X, ym = 2Y, xf = yf = 128 (ago_haf_cpu_geometric.cpp:2000-2045 with
        // xscale = yscale = 2, offsets 0.5): two-stage rounded averages; tail columns single rounding
        const int X = tx0 / 2 + lane;
        const int body = P.dw & ~7;
        if (X < P.dw) {
            for (int oy = wave; oy < TH / 2; oy += NW) {
                const int Yo = ty0 / 2 + oy;
                if (Yo >= P.dh) break;
                const int sy = 2 * oy + HY, sx = 2 * lane + HY;
                const int p00 = s.Y[sy][sx], p01 = s.Y[sy][sx + 1], p10 = s.Y[sy + 1][sx], p11 = s.Y[sy + 1][sx + 1];
                int v;
                if (X < body) {
                    const int t0 = (128 * p00 + 128 * p01 + 0x80) >> 8, t1 = (128 * p10 + 128 * p11 + 0x80) >> 8;
                    v = (128 * t0 + 128 * t1 + 0x80) >> 8;
                } else {
                    v = (128 * 128 * (p00 + p01 + p10 + p11) + 0x8000) >> 16;
                }
                gmem(F.scaled)[(long)Yo * F.scaledStride + X] = (uint8_t)v;
            }
        }
        // ScaleGaussianHalf_5x5 level 1 (ago_haf_cpu_pyramid.cpp:94-305): centre (2X+ox, 2Y+oy)
        if (P.levels > 1 && X < P.lvlW[1]) {
            const int oxp = (w & 1) ? 0 : 1, oyp = (h & 1) ? 0 : 1;
            const int l1h = P.lvlH[1];
            for (int oy = wave; oy < TH / 2; oy += NW) {
                const int Yo = ty0 / 2 + oy;
                if (Yo > l1h - 2) break;
                if (Yo < 1) continue;
                const int cy = 2 * oy + oyp + HY, cx = 2 * lane + oxp + HY;
                uint32_t acc = 0;
#pragma unroll
                for (int jj = 0; jj < 5; jj++) {
                    const uint8_t *rr = &s.Y[cy - 2 + jj][cx - 2];
                    const uint32_t hr = rr[0] + 4u * rr[1] + 6u * rr[2] + 4u * rr[3] + rr[4];
                    acc += (jj == 0 || jj == 4 ? 1u : jj == 2 ? 6u : 4u) * hr;
                }
                gmem(F.lvl[1])[(long)Yo * F.lvlStride[1] + X] = (uint8_t)((acc & 0xffffu) >> 8);
            }
        }
    }

    // 6. tile-local hysteresis.  (a) One wave, lane = row, 64-bit row masks, dilate-and-mask until
    //    nothing changes, two problems side by side: lanes 0..31 grow the edge pixels that a strong
    //    pixel reaches inside the tile (the union of the 8-connected edge components that hold a
    //    strong pixel), lanes 32..63 grow those that a tile-border edge pixel reaches (the components
    //    that touch the tile border).  Edge pixels in neither are weak components inside the tile:
    //    they settle to 0 here.  Border-touching components without a strong pixel are the only ones
    //    left undecided.
    static_assert(TH <= 32 && TW <= 63, "row masks: one lane per row in each half-wave");
    __syncthreads();
    if (wave == 0) {
        const int r = lane & 31;
        const uint64_t e = r < TH ? s.Em[r] : 0ull;
        uint64_t m;
        if (lane < 32) m = r < TH ? s.Sm[r] : 0ull;
        else m = (r == 0 || r == th - 1) ? e : (e & (1ull | (1ull << (tw - 1))));
        while (true) {
            uint64_t up = __shfl(m, lane > 0 ? lane - 1 : 0), dn = __shfl(m, lane < 63 ? lane + 1 : 63);
            if (r == 0) up = 0ull;
            if (r >= TH - 1) dn = 0ull;
            uint64_t v = m | up | dn;
            v |= (v << 1) | (v >> 1);
            const uint64_t n = (v & e) | m;
            const bool changed = n != m;
            m = n;
            if (!__any(changed)) break;
        }
        const uint64_t other = __shfl(m, lane ^ 32);
        if (r < TH) {
            if (lane < 32) s.Rm[r] = m;
            else s.Bm[r] = m & ~other;
        }
    }
    __syncthreads();
    //    (b) union-find over the undecided pixels (8-connectivity in the tile).  Every 8-neighbour
    //    of an undecided pixel that is an edge pixel lies in the same component, so the row masks
    //    answer the neighbour tests and only undecided pixels are labelled.
    for (int q = 0; q < TH / NW; q++) {
        const int r = wave + q * NW;
        if ((s.Bm[r] >> lane) & 1ull) s.L[r * TP + lane] = r * TP + lane;
    }
    __syncthreads();
    for (int q = 0; q < TH / NW; q++) {
        const int r = wave + q * NW;
        const uint64_t b = s.Bm[r];
        if (!b) continue;
        if (!((b >> lane) & 1ull)) continue;
        const int p = r * TP + lane;
        if ((b >> (lane + 1)) & 1ull) lunion(s.L, p, p + 1);
        if (r + 1 < TH) {
            const uint64_t bd = s.Bm[r + 1];
            if (lane > 0 && ((bd >> (lane - 1)) & 1ull)) lunion(s.L, p, p + TP - 1);
            if ((bd >> lane) & 1ull) lunion(s.L, p, p + TP);
            if ((bd >> (lane + 1)) & 1ull) lunion(s.L, p, p + TP + 1);
        }
    }
    __syncthreads();
    //    (c) settle: reached -> 255; other edge pixels -> 0; undecided -> 127 and listed with their
    //    tile-local root for k_border / k_resolve, under global id 1 + tile*TS + root.  Every reached
    //    border pixel maps to the single strong node 0 (the smallest id, so after atomicMin linking a
    //    set is strong iff its root is 0).
    for (int q = 0; q < TH / NW; q++) {
        const int r = wave + q * NW;
        const int y = ty0 + r, x = tx0 + lane;
        if (lane >= TW || !(INNER || (y >= 1 && y <= h - 2 && x >= 1 && x <= w - 2))) continue;
        uint8_t o = 0;
        if ((s.Rm[r] >> lane) & 1ull) {
            o = 255;
        } else if ((s.Bm[r] >> lane) & 1ull) {
            const int p = r * TP + lane;
            const int root = lfind(s.L, p);
            o = 127;
            if (root == p) gmem(F.parent)[1u + (uint32_t)tile * TS + (uint32_t)p] = 1u + (uint32_t)tile * TS + (uint32_t)p;
            const uint32_t i = atomicAdd(&s.unres, 1u);
            gmem(F.ulist)[(size_t)tile * TS + i] = (uint32_t)p | ((uint32_t)root << 16);
        }
        gmem(F.edges)[(long)y * F.edgesStride + x] = o;
    }
    // border tables: top row, bottom row, left column, right column of the tile (actual extent)
    gptr_t<int32_t> bt = gmem(F.border) + (long)tile * 2 * (TW + TH);
    for (int n = t; n < 2 * (TW + TH); n += FT) {
        int lx, ly;
        if (n < TW) { lx = n; ly = 0; }
        else if (n < 2 * TW) { lx = n - TW; ly = th - 1; }
        else if (n < 2 * TW + TH) { lx = 0; ly = n - 2 * TW; }
        else { lx = tw - 1; ly = n - 2 * TW - TH; }
        int32_t g = -1;
        if (lx < tw && ly < th) {
            if ((s.Rm[ly] >> lx) & 1ull) g = 0;
            else if ((s.Bm[ly] >> lx) & 1ull) g = (int32_t)(1u + (uint32_t)tile * TS + (uint32_t)lfind(s.L, ly * TP + lx));
        }
        bt[n] = g;
    }
    __syncthreads();
    if (t == 0) gmem(F.ucount)[tile] = s.unres;
}

__global__ void __launch_bounds__(FT)
k_front(const FrameDesc *__restrict__ frames, GraphParams P)
{
    __shared__ FrontShared s;
    const FrameDesc F = frames[blockIdx.z];
    const int tx0 = (int)blockIdx.x * TW, ty0 = (int)blockIdx.y * TH;
    const int tile = (int)blockIdx.y * P.tilesX + (int)blockIdx.x;
    if (threadIdx.x == 0) { s.cnt = 0; s.unres = 0; }
    const bool inner = tx0 >= 4 && tx0 + TW + 4 <= P.w && ty0 >= 4 && ty0 + TH + 4 <= P.h;
    if (inner) front_tile<true>(s, F, P, tx0, ty0, tile);
    else front_tile<false>(s, F, P, tx0, ty0, tile);
}

// union the border components of each tile with those of its right, lower and diagonal neighbours;
// one wave per tile (lane = row of the right column, then column of the bottom row), four tiles per
// workgroup
__global__ void __launch_bounds__(256)
k_border(const FrameDesc *__restrict__ frames, GraphParams P)
{
    static_assert(TW <= 64 && TH <= 64, "one lane per border pixel");
    const FrameDesc &F = frames[blockIdx.z];
    const int lane = (int)threadIdx.x & 63;
    const int tile = (int)blockIdx.x * 4 + ((int)threadIdx.x >> 6);
    if (tile >= P.tilesX * P.tilesY) return;
    const int tX = tile % P.tilesX, tYi = tile / P.tilesX;
    const int tw = min(TW, P.w - tX * TW), th = min(TH, P.h - tYi * TH);
    gptr_t<const int32_t> tables = gmem((const int32_t *)F.border);
    gptr_t<const int32_t> bt = tables + (long)tile * 2 * (TW + TH);
    gptr_t<uint32_t> par = gmem(F.parent);
    // right neighbour: my right column (rows 0..th-1) against its left column, |dr| <= 1
    if (tX + 1 < P.tilesX && lane < th) {
        gptr_t<const int32_t> nb = tables + (long)(tile + 1) * 2 * (TW + TH);
        const int32_t a = bt[2 * TW + TH + lane];
        if (a >= 0) {
            for (int d = -1; d <= 1; d++) {
                const int rr = lane + d;
                if (rr < 0 || rr >= th) continue;
                const int32_t b = nb[2 * TW + rr];
                if (b >= 0 && b != a) gunion(par, (uint32_t)a, (uint32_t)b);
            }
        }
    }
    // lower neighbour: my bottom row against its top row, |dc| <= 1 (the tile below has the same width)
    if (tYi + 1 < P.tilesY) {
        gptr_t<const int32_t> nb = tables + (long)(tile + P.tilesX) * 2 * (TW + TH);
        if (lane < tw) {
            const int32_t a = bt[TW + lane];
            if (a >= 0) {
                for (int d = -1; d <= 1; d++) {
                    const int cc = lane + d;
                    if (cc < 0 || cc >= tw) continue;
                    const int32_t b = nb[cc];
                    if (b >= 0 && b != a) gunion(par, (uint32_t)a, (uint32_t)b);
                }
            }
        }
        if (lane == 0) {
            // diagonals: bottom-right corner <-> lower-right tile's top-left,
            // bottom-left corner <-> lower-left tile's top-right (a full-width tile)
            if (tX + 1 < P.tilesX) {
                const int32_t a = bt[TW + tw - 1];
                gptr_t<const int32_t> nd = tables + (long)(tile + P.tilesX + 1) * 2 * (TW + TH);
                if (a >= 0 && nd[0] >= 0 && nd[0] != a) gunion(par, (uint32_t)a, (uint32_t)nd[0]);
            }
            if (tX > 0) {
                const int32_t a = bt[TW + 0];
                gptr_t<const int32_t> nd = tables + (long)(tile + P.tilesX - 1) * 2 * (TW + TH);
                if (a >= 0 && nd[TW - 1] >= 0 && nd[TW - 1] != a) gunion(par, (uint32_t)a, (uint32_t)nd[TW - 1]);
            }
        }
    }
}

// the undecided pixels k_front listed take the state of their component's global root (strong iff
// the root is node 0); one wave per tile, four tiles per workgroup
__global__ void __launch_bounds__(256)
k_resolve(const FrameDesc *__restrict__ frames, GraphParams P)
{
    const FrameDesc &F = frames[blockIdx.z];
    const int lane = (int)threadIdx.x & 63;
    const int tile = (int)blockIdx.x * 4 + ((int)threadIdx.x >> 6);
    if (tile >= P.tilesX * P.tilesY) return;
    const uint32_t n = gmem(F.ucount)[tile];
    if (n == 0) return;
    const int tx0 = (tile % P.tilesX) * TW, ty0 = (tile / P.tilesX) * TH;
    gptr_t<const uint32_t> list = gmem((const uint32_t *)F.ulist) + (size_t)tile * TS;
    for (uint32_t i = (uint32_t)lane; i < n; i += 64) {
        const uint32_t e = list[i];
        const uint32_t p = e & 0xffffu, root = e >> 16;
        const uint32_t g = gfind(gmem(F.parent), 1u + (uint32_t)tile * TS + root);
        gmem(F.edges)[(long)(ty0 + (int)(p / TP)) * F.edgesStride + tx0 + (int)(p % TP)] = (g == 0) ? 255 : 0;
    }
}

// pyramid level `lvl` (>= 2) from level lvl-1, all frames
__global__ void __launch_bounds__(256)
k_half_batched(const FrameDesc *__restrict__ frames, GraphParams P, int lvl)
{
    const FrameDesc &F = frames[blockIdx.z];
    const uint32_t x0 = (blockIdx.x * 64 + threadIdx.x) * 8;
    const int y0 = 1 + (int)(blockIdx.y * 4 + threadIdx.y) * PYR_ROWS;
    half_strip((uint32_t)P.lvlW[lvl], (uint32_t)P.lvlH[lvl], F.lvl[lvl], F.lvlStride[lvl], (uint32_t)P.lvlW[lvl - 1],
               (uint32_t)P.lvlH[lvl - 1], F.lvl[lvl - 1], F.lvlStride[lvl - 1], x0, y0);
}

// Harris sort and pick, one workgroup per frame; clears the candidate count for the next batch
__global__ void __launch_bounds__(PICK_THREADS)
k_pick_batched(const FrameDesc *__restrict__ frames, GraphParams P)
{
    const FrameDesc &F = frames[blockIdx.x];
    const PickWork W = *reinterpret_cast<const PickWork *>(F.pickWs);
    harris_pick_block((const XYS *)F.cand, F.candCount, true, P.candCap, P.dist, (uint32_t)P.w, (uint32_t)P.h,
                      (Keypoint *)F.corners, P.capacity, F.numItems, F.count, W);
}

__global__ void k_fill_bytes(uint32_t *p, uint32_t v, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

// ---------------------------------------------------------------- host side -----------------------
struct GraphLayout {
    size_t descs, frameBytes, border, parent, ulist, ucount, cand, candCount, pickHdr, pickWs, total;
    uint32_t tiles, candCap;
};

inline GraphLayout graph_layout(const mvx_vision_params_t &p, uint32_t maxFrames)
{
    GraphLayout L;
    const uint32_t tx = (p.width + TW - 1) / TW, ty = (p.height + TH - 1) / TH;
    L.tiles = tx * ty;
    L.candCap = std::max(1u, ((p.width - 1) / 2) * ((p.height - 1) / 2));   // NMS maxima are never 8-adjacent
    L.descs = align256(sizeof(FrameDesc) * (size_t)maxFrames);
    size_t o = 0;
    L.border = o;    o = align256(o + (size_t)L.tiles * 2 * (TW + TH) * 4);
    L.parent = o;    o = align256(o + ((size_t)L.tiles * TS + 1) * 4);
    L.ulist = o;     o = align256(o + (size_t)L.tiles * TS * 4);
    L.ucount = o;    o = align256(o + (size_t)L.tiles * 4);
    L.cand = o;      o = align256(o + (size_t)L.candCap * sizeof(XYS));
    L.candCount = o; o = align256(o + 4);
    L.pickHdr = o;   o = align256(o + sizeof(PickWork));
    L.pickWs = o;    o = align256(o + pick_layout(L.candCap).total);
    L.frameBytes = o;
    L.total = L.descs + L.frameBytes * maxFrames;
    return L;
}

inline bool params_ok(const mvx_vision_params_t &p)
{
    if (!(p.width >= 16 && p.height >= 16 && (p.width % 4) == 0 && p.width <= 32767 && p.height <= 32767 &&
          (p.height % 2) == 0 && p.scaledWidth == p.width / 2 && p.scaledHeight == p.height / 2 && p.levels >= 1 &&
          p.levels <= 8 && p.harrisMinDistance > 0.0f))
        return false;
    // every pyramid step must be one the reference accepts (source >= 5, destination >= 3)
    uint32_t lw = p.width, lh = p.height;
    for (uint32_t l = 1; l < p.levels; l++) {
        const uint32_t nw = (lw + 1) / 2, nh = (lh + 1) / 2;
        if (lw < 5 || lh < 5 || nw < 3 || nh < 3) return false;
        lw = nw; lh = nh;
    }
    return true;
}

inline GraphParams graph_params(const mvx_vision_params_t &p, const GraphLayout &L)
{
    GraphParams P;
    P.w = (int)p.width; P.h = (int)p.height;
    P.yStride = (int)((p.width + 15) & ~15u);
    P.tilesX = (int)((p.width + TW - 1) / TW); P.tilesY = (int)((p.height + TH - 1) / TH);
    P.dw = (int)p.scaledWidth; P.dh = (int)p.scaledHeight;
    P.levels = (int)p.levels;
    uint32_t lw = p.width, lh = p.height;
    for (int l = 0; l < 8; l++) { P.lvlW[l] = (int)lw; P.lvlH[l] = (int)lh; lw = (lw + 1) / 2; lh = (lh + 1) / 2; }
    P.lower = p.cannyLower; P.upper = p.cannyUpper;
    P.k = p.harrisK; P.thr = p.harrisThreshold; P.dist = p.harrisMinDistance;
    // normFactor of ago_kernel_api.cpp:16941-16942 for gradient size 3, in float, left to right
    float nf = 255.0f * (float)(1 << 2) * 3;
    P.norm = nf * nf * nf * nf;
    P.normRcp = recip_rn(P.norm);
    P.capacity = p.cornerCapacity;
    P.candCap = L.candCap;
    return P;
}

inline bool plane_ok_v(const void *ptr, uint32_t w, uint32_t stride, uint32_t bpp)
{
    return ptr && aligned16(ptr) && (stride % 16u) == 0 && stride >= ((w * bpp + 15u) & ~15u);
}

}  // namespace mvx

using namespace mvx;

extern "C" int mvx_VisionGraph_supported(const mvx_vision_params_t *p) { return (p && params_ok(*p)) ? 1 : 0; }

extern "C" size_t mvx_VisionGraph_workspace_size(const mvx_vision_params_t *p, uint32_t maxFrames)
{
    if (!p || !params_ok(*p) || maxFrames == 0) return 0;
    return graph_layout(*p, maxFrames).total;
}

extern "C" int mvx_VisionGraph_prepare(mvx_stream_t stream, const mvx_vision_params_t *p, const mvx_vision_frame_t *frames,
                                       uint32_t nframes, void *workspace, size_t workspaceSize)
{
    if (!p || !params_ok(*p) || !frames || nframes == 0 || !workspace) return MVX_ERROR_INVALID_PARAMETERS;
    const GraphLayout L = graph_layout(*p, nframes);
    if (workspaceSize < L.total || !aligned16(workspace)) return MVX_ERROR_INVALID_PARAMETERS;
    std::vector<FrameDesc> D(nframes);
    uint8_t *base = (uint8_t *)workspace + L.descs;
    uint32_t lw = p->width, lh = p->height;
    for (uint32_t f = 0; f < nframes; f++) {
        const mvx_vision_frame_t &in = frames[f];
        FrameDesc &d = D[f];
        memset(&d, 0, sizeof(d));
        if (!plane_ok_v(in.rgb, p->width, in.rgbStride, 3) || !plane_ok_v(in.scaled, p->scaledWidth, in.scaledStride, 1) ||
            !plane_ok_v(in.edges, p->width, in.edgesStride, 1) || !in.corners || !in.numItems || !in.numCorners)
            return MVX_ERROR_INVALID_PARAMETERS;
        lw = p->width; lh = p->height;
        for (uint32_t l = 0; l < p->levels; l++) {
            if (!plane_ok_v(in.pyramid[l], lw, in.pyramidStride[l], 1)) return MVX_ERROR_INVALID_PARAMETERS;
            d.lvl[l] = in.pyramid[l];
            d.lvlStride[l] = in.pyramidStride[l];
            lw = (lw + 1) / 2; lh = (lh + 1) / 2;
        }
        d.rgb = in.rgb; d.rgbStride = in.rgbStride;
        d.scaled = in.scaled; d.scaledStride = in.scaledStride;
        d.edges = in.edges; d.edgesStride = in.edgesStride;
        d.corners = in.corners; d.numItems = in.numItems; d.count = in.numCorners;
        uint8_t *fb = base + (size_t)f * L.frameBytes;
        d.border = (int32_t *)(fb + L.border);
        d.parent = (uint32_t *)(fb + L.parent);
        d.ulist = (uint32_t *)(fb + L.ulist);
        d.ucount = (uint32_t *)(fb + L.ucount);
        d.cand = fb + L.cand;
        d.candCount = (uint32_t *)(fb + L.candCount);
        d.pickWs = fb + L.pickHdr;
        // the pick workspace header holds the device pointers of the frame's pick scratch
        const PickWork W = pick_work(fb + L.pickWs, L.candCap);
        if (hipMemcpy(fb + L.pickHdr, &W, sizeof(W), hipMemcpyHostToDevice) != hipSuccess) return MVX_FAILURE;
        const PickLayout PL = pick_layout(L.candCap);
        hipLaunchKernelGGL(k_fill_bytes, dim3(256), dim3(256), 0, stream_of(stream), (uint32_t *)(fb + L.pickWs + PL.hkey),
                           HASH_EMPTY, (size_t)PL.hsize);
        hipLaunchKernelGGL(k_fill_bytes, dim3(1), dim3(64), 0, stream_of(stream), d.candCount, 0u, (size_t)1);
        // union-find slot 0 is the strong node every strong tile-border pixel maps to: its own root
        hipLaunchKernelGGL(k_fill_bytes, dim3(1), dim3(64), 0, stream_of(stream), d.parent, 0u, (size_t)1);
    }
    if (hipMemcpy(workspace, D.data(), sizeof(FrameDesc) * nframes, hipMemcpyHostToDevice) != hipSuccess) return MVX_FAILURE;
    MVX_LAUNCH_CHECK();
    return MVX_SUCCESS;
}

extern "C" int mvx_VisionGraph_run_stage(mvx_stream_t stream, const mvx_vision_params_t *p, uint32_t nframes,
                                         void *workspace, size_t workspaceSize, int stage)
{
    if (!p || !params_ok(*p) || nframes == 0 || !workspace || stage < 0 || stage >= MVX_VISION_STAGES)
        return MVX_ERROR_INVALID_PARAMETERS;
    const GraphLayout L = graph_layout(*p, nframes);
    if (workspaceSize < L.total) return MVX_ERROR_INVALID_PARAMETERS;
    const GraphParams P = graph_params(*p, L);
    const FrameDesc *D = (const FrameDesc *)workspace;
    hipStream_t s = stream_of(stream);
    switch (stage) {
    case 0:
        hipLaunchKernelGGL(k_front, dim3(P.tilesX, P.tilesY, nframes), dim3(FT), 0, s, D, P);
        break;
    case 1:
        hipLaunchKernelGGL(k_border, dim3((P.tilesX * P.tilesY + 3) / 4, 1, nframes), dim3(256), 0, s, D, P);
        break;
    case 2:
        hipLaunchKernelGGL(k_resolve, dim3((P.tilesX * P.tilesY + 3) / 4, 1, nframes), dim3(256), 0, s, D, P);
        break;
    case 3:
        for (int l = 2; l < P.levels; l++) {
            const uint32_t chunks = ((uint32_t)P.lvlW[l] + 7) / 8;
            const uint32_t strips = ((uint32_t)P.lvlH[l] - 2 + PYR_ROWS - 1) / PYR_ROWS;
            hipLaunchKernelGGL(k_half_batched, dim3((chunks + 63) / 64, (strips + 3) / 4, nframes), dim3(64, 4), 0, s, D,
                               P, l);
        }
        break;
    case 4:
        hipLaunchKernelGGL(k_pick_batched, dim3(nframes), dim3(PICK_THREADS), 0, s, D, P);
        break;
    }
    MVX_LAUNCH_CHECK();
    return MVX_SUCCESS;
}

namespace mvx {

// The stream the pyramid levels 2.. and the Harris pick run on beside the hysteresis stages, with the
// fork and join events: one set per device, default priority, created on first use (outside any
// capture, normally) and kept until the process ends.  The mutex covers the enqueue sequence, so
// calls from several host threads reuse the events safely.
struct SideLane {
    hipStream_t stream = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    bool ok = false, tried = false;
};
static std::mutex g_sideMu;
static SideLane g_side[64];

static SideLane *side_lane()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return nullptr;
    SideLane &L = g_side[dev];
    if (!L.tried) {
        L.tried = true;
        L.ok = hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking) == hipSuccess &&
               hipEventCreateWithFlags(&L.fork, hipEventDisableTiming) == hipSuccess &&
               hipEventCreateWithFlags(&L.join, hipEventDisableTiming) == hipSuccess;
    }
    return L.ok ? &L : nullptr;
}

}  // namespace mvx

// Stage 0 on the caller's stream; then stages 3 and 4 (which need only stage 0) on the side stream
// while stages 1 and 2 run on the caller's stream; the caller's stream waits for the side stream at
// the end, so the call orders like a single-stream sequence.  Fork and join are event waits, which a
// hipGraph capture of the caller's stream records as dependencies.
extern "C" int mvx_VisionGraph_run(mvx_stream_t stream, const mvx_vision_params_t *p, uint32_t nframes, void *workspace,
                                   size_t workspaceSize)
{
    int rc = mvx_VisionGraph_run_stage(stream, p, nframes, workspace, workspaceSize, 0);
    if (rc) return rc;
    std::lock_guard<std::mutex> lock(g_sideMu);
    SideLane *L = side_lane();
    if (!L) {
        for (int st = 1; st < MVX_VISION_STAGES && !rc; st++)
            rc = mvx_VisionGraph_run_stage(stream, p, nframes, workspace, workspaceSize, st);
        return rc;
    }
    hipStream_t main = stream_of(stream);
    const mvx_stream_t side = reinterpret_cast<mvx_stream_t>(L->stream);
    if (hipEventRecord(L->fork, main) != hipSuccess || hipStreamWaitEvent(L->stream, L->fork, 0) != hipSuccess)
        return MVX_FAILURE;
    if ((rc = mvx_VisionGraph_run_stage(side, p, nframes, workspace, workspaceSize, 3))) return rc;
    if ((rc = mvx_VisionGraph_run_stage(side, p, nframes, workspace, workspaceSize, 4))) return rc;
    if (hipEventRecord(L->join, L->stream) != hipSuccess) return MVX_FAILURE;
    if ((rc = mvx_VisionGraph_run_stage(stream, p, nframes, workspace, workspaceSize, 1))) return rc;
    if ((rc = mvx_VisionGraph_run_stage(stream, p, nframes, workspace, workspaceSize, 2))) return rc;
    return hipStreamWaitEvent(main, L->join, 0) == hipSuccess ? MVX_SUCCESS : MVX_FAILURE;
}
