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
//   k_front     one 58x32 luma tile (+3 halo) per wave, four independent waves per workgroup.  Every
//               window row of the tile (64 luma, 62 gradient, 60 score columns) fits the 64 lanes, so
//               lane = column and the wave streams the tile's rows top to bottom with every quantity in
//               rolling register windows (column neighbours through DPP lane shifts): RGB -> Y, level 0,
//               the 2:1 scaled image, level 1, Canny gradients/NMS/threshold, Harris products / exact
//               integer window sums / score / NMS; then tile-local hysteresis on 64-bit row masks
//               (lane = row) with the border-touching weak components labelled by flood fill.  No
//               barriers; LDS holds only the wave's Harris candidates.  Only graph outputs and small
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
#include "rowmask.hpp"
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
constexpr int TNB = 2 * (TW + TH);               // union-find slots per tile: one per border pixel, at most
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
    uint32_t *parent;     // 1 + tiles * TNB union-find slots: 0 the strong node, then TNB per tile
    uint32_t *ulist;      // tiles x TS: a tile's undecided pixels, p | the component's number k << 16
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

MVX_DEV float lane_next_f(float v) { return __int_as_float(lane_next(__float_as_int(v))); }

constexpr int FTW = 4;                                     // tiles (one wave each) per workgroup of k_front
constexpr int RES_TILES = 4;                               // tiles per wave of k_resolve
constexpr int CAND_MAX = ((TW + 1) / 2) * ((TH + 1) / 2);  // Harris NMS maxima of a tile are never 8-adjacent

// One wave streams one tile: luma window rows k = 0..YH-1 (image row ty0-3+k, lane = column tx0-3+lane)
// top to bottom, every later quantity from rolling register windows of the rows it needs, column
// neighbours by DPP lane shifts.  At step k:
//   luma row k -> pyramid level 0 (row k-3), the 2:1 scaled image and level 1 when their rows complete;
//   gradient row g = k-2 (image row ty0-2+g, column tx0-2+lane) -> HG3 products and 3-column sums;
//   Canny magnitude|sector row k-3 and Harris score row k-4 (image row ty0-1+m, column tx0-1+lane);
//   NMS of tile row r = k-6 (image row ty0+r, column tx0+lane): Canny double threshold into 64-bit row
//   masks (edge, strong; lane r holds row r), Harris maxima into the per-wave candidate buffer.
// Then the tile's hysteresis on the row masks and the settle pass.
template <bool INNER>
MVX_DEV void front_tile(uint64_t *cand, const FrameDesc &F, const GraphParams &P, int tx0, int ty0, int tile)
{
    const int lane = (int)threadIdx.x & 63;
    const int w = P.w, h = P.h;
    const int tw = INNER ? TW : min(TW, w - tx0), th = INNER ? TH : min(TH, h - ty0);
    const int x = tx0 - 1 + lane;       // Canny magnitude / Harris score column
    const int ox = tx0 + lane;          // tile output column

    // inner tiles read a pixel's R, G, B (and the next pixel's R) with one dword load; the window's
    // last column is at most w-2, so the dword stays inside the row
    const __amdgpu_buffer_rsrc_t rgbR = plane_rsrc(F.rgb), l0R = plane_rsrc(F.lvl[0]), scR = plane_rsrc(F.scaled);
    const __amdgpu_buffer_rsrc_t l1R = plane_rsrc(F.lvl[1]), edR = plane_rsrc(F.edges);
    const int rgbCol = 3 * (tx0 - HY + lane);
    auto fetch = [&](int k) -> uint32_t {
        return __builtin_amdgcn_raw_buffer_load_b32(rgbR, rgbCol, (int)((uint32_t)(ty0 - HY + k) * F.rgbStride), 0);
    };
    const int body = P.dw & ~7;
    const int oxp = (w & 1) ? 0 : 1, oyp = (h & 1) ? 0 : 1;
    // Harris scores compare as int32 keys: a non-negative float's bits order like the float; with a
    // negative threshold negative scores occur, and the key flips their magnitude bits (an involution).
    // Scores are never -0 (mc is never -0), so equal keys are equal scores.
    const bool negKeys = P.thr < 0.0f;
    const int U = (int)P.upper, Lo = (int)P.lower;
    // output columns of this lane: level 0 (luma lane), 2:1 scaled (lane 2X'+3), level 1 (lane 2X'+oxp+1)
    const int sp = lane - HY, lp1 = lane - oxp - 1;
    const bool scOut = sp >= 0 && !(sp & 1) && (sp >> 1) < TW / 2 && tx0 / 2 + (sp >> 1) < P.dw;
    const bool l1Out = lp1 >= 0 && !(lp1 & 1) && (lp1 >> 1) < TW / 2 && tx0 / 2 + (lp1 >> 1) < P.lvlW[1];
    const bool scBody = tx0 / 2 + (sp >> 1) < body;

    uint32_t emLo = 0, emHi = 0, smLo = 0, smHi = 0;      // lane r: edge / strong mask of tile row r
    uint32_t nc = 0;                                      // candidates in the wave's buffer
    uint32_t packedPrev = 0;                              // Canny magnitude|sector of gradient row k-3, its own column
    int Vb = 0;                                           // Harris score key row k-5

    // One step of the stream (luma row k).  Windows deeper than one row are passed in age order and
    // the step writes its new row over the oldest one, so the caller rotates the arguments instead of
    // the values (no register copies): yA..yD = luma rows k-4..k-1 (yA <- row k); dO, dN / sO, sN = rows
    // k-2, k-1 (dO <- row k); xx0.. = HG3 sums of gradient rows g-2, g-1 (<- row g); MO, MN = Canny
    // rows k-6, k-5 (MO <- row k-4); HO, HN = score key 3-column maxima rows k-6, k-5 (HO <- row k-4);
    // pO, pN = RGB of rows k, k+1 (pO <- row k+2).
    auto step = [&](const int k, uint32_t &pO, uint32_t &pN, int &yA, int &yB, int &yC, int &yD, int &dO, int &dN,
                    int &sO, int &sN, int &xx0, int &xx1, int &xy0, int &xy1, int &yy0, int &yy1, uint32_t &MO,
                    uint32_t &MN, int &HO, int &HN) {
        // ---- luma row k (ColorConvert Y of the RGB pixel, or the virtual plane's linear-address
        //      value); the float sum lies in [0, 256), so the truncation needs no clamp
        int yk;
        if (INNER) {
            const uint32_t cur = pO;
            if (k + 2 < YH) pO = fetch(k + 2);
            float f = (float)(cur & 0xffu) * 0.2126f;
            f = f + (float)((cur >> 8) & 0xffu) * 0.7152f;
            f = f + (float)((cur >> 16) & 0xffu) * 0.0722f;
            yk = (int)f;
        } else {
            yk = (int)yflat(F, P, ty0 - HY + k, tx0 - HY + lane);
        }
        if (k >= HY && k < HY + th && lane >= HY && lane < HY + tw)
            __builtin_amdgcn_raw_buffer_store_b8((uint8_t)yk, l0R, tx0 - HY + lane, (int)((uint32_t)(ty0 + k - HY) * F.lvlStride[0]), 0);
        // binomial column sums through one-lane shifts: t = Y[l]+Y[l+1], s = Y[l]+2Y[l+1]+Y[l+2], d = Y[l+2]-Y[l]
        const int tk = yk + lane_next(yk);
        const int sk = tk + lane_next(tk), dk = lane_next(tk) - tk;

        // ---- ScaleImage 2:1 (ago_haf_cpu_geometric.cpp:2000-2045 with xscale = yscale = 2, offsets 0.5):
        //      output row oy from luma rows 2oy+3 (k-1) and 2oy+4 (k); lane 2X'+3 holds output column X'
        if (k >= 4 && !(k & 1) && k <= HY + TH - 1) {
            const int Yo = ty0 / 2 + ((k - 4) >> 1);
            const int top = yD + lane_next(yD);                 // p00 + p01; p10 + p11 = tk
            if (Yo < P.dh && scOut) {
                const int v = scBody ? ((((top + 1) >> 1) + ((tk + 1) >> 1) + 1) >> 1)   // (128 a + 128 b + 0x80) >> 8, twice
                                     : ((top + tk + 2) >> 2);                             // (128*128 sum + 0x8000) >> 16
                __builtin_amdgcn_raw_buffer_store_b8((uint8_t)v, scR, tx0 / 2 + (sp >> 1), (int)((uint32_t)Yo * F.scaledStride), 0);
            }
        }
        // ---- ScaleGaussianHalf_5x5 level 1 (ago_haf_cpu_pyramid.cpp:94-305): centre (2X+oxp, 2Y+oyp),
        //      luma rows k-4..k; the horizontal 1 4 6 4 1 is four "v + v[l+1]" steps, so lane 2X'+oxp+1
        //      holds the sum over columns lane..lane+4 (centre lane+2)
        if (P.levels > 1 && k >= 5 + oyp && !((k - 5 - oyp) & 1)) {
            const int oy = (k - 5 - oyp) >> 1, Yo = ty0 / 2 + oy;
            if (oy < TH / 2 && Yo >= 1 && Yo <= P.lvlH[1] - 2) {
                int acc = (yA + yk) + 4 * (yB + yD) + 6 * yC;
                acc = acc + lane_next(acc);
                acc = acc + lane_next(acc);
                acc = acc + lane_next(acc);
                acc = acc + lane_next(acc);
                if (l1Out)
                    __builtin_amdgcn_raw_buffer_store_b8((uint8_t)(((uint32_t)acc & 0xffffu) >> 8), l1R, tx0 / 2 + (lp1 >> 1),
                                                         (int)((uint32_t)Yo * F.lvlStride[1]), 0);
            }
        }

        // ---- Sobel gradient row g = k-2 (linear-address luma; |gx|, |gy| <= 1020, so the reference's S16
        //      values are these), HG3 products and 3-column sums.  HG3 rows 0 and h-1 are never written
        //      (zero).  The products are integers whose 9-sums stay below 2^24, so int32 sums equal the
        //      reference's float sums exactly.
        const int g = k - 2;
        const int gx = dO + 2 * dN + dk, gy = sk - sO;
        int pxx = 0, pxy = 0, pyy = 0;
        if (INNER || (ty0 - 2 + g >= 1 && ty0 - 2 + g <= h - 2)) { pxx = __mul24(gx, gx); pxy = __mul24(gx, gy); pyy = __mul24(gy, gy); }
        asm("" : "+v"(pxx), "+v"(pxy), "+v"(pyy));   // keep the window sums as shifted adds (not mads)
        const int hxx = pxx + lane_next(pxx + lane_next(pxx));
        const int hxy = pxy + lane_next(pxy + lane_next(pxy));
        const int hyy = pyy + lane_next(pyy + lane_next(pyy));

        // ---- Harris score row k-4: Vc = mc / norm > T ? mc / norm : 0, the quotient exactly the IEEE one;
        //      kept as its compare key, with the maximum of the key over columns lane..lane+2
        int Vc = 0;
        {
            const int yv = ty0 - 1 + (k - 4);
            float v = 0.0f;
            if (k >= 4 && (INNER || (x >= 1 && x <= w - 2 && yv >= 1 && yv <= h - 2))) {
                const float gx2 = (float)(xx0 + xx1 + hxx);
                const float gxy2 = (float)(xy0 + xy1 + hxy);
                const float gy2 = (float)(yy0 + yy1 + hyy);
                const float tr = gx2 + gy2;
                const float det = (gx2 * gy2) - (gxy2 * gxy2);
                const float mc = det - (P.k * tr * tr);
                const float q = div_by_const_rn(mc, P.norm, P.normRcp);
                v = (q > P.thr) ? q : 0.0f;
            }
            Vc = __float_as_int(v);
            if (negKeys) Vc ^= (Vc >> 31) & 0x7fffffff;
        }
        int Hc = max(Vc, lane_next(Vc));
        Hc = max(Vc, lane_next(Hc));

        // ---- Canny magnitude|sector row m = k-4 from gradient row k-3 (packed at its own column, moved
        //      to the magnitude column: lane+1 -> lane)
        uint32_t Mc = 0;
        {
            const int m = k - 4, y = ty0 - 1 + m;
            const uint32_t shifted = (uint32_t)lane_next((int)packedPrev);
            if (k >= 4 && (INNER || (y >= 1 && y <= h - 2 && x >= 1 && x < w))) {
                Mc = shifted;
            } else if (!INNER && k >= 4 && y >= 1 && y <= h - 2 && tx0 == 0) {
                // CannySobel column 0 (ago_haf_cpu_canny.cpp:83-116 with the staged rows r0[-1] = 0 and
                // r1[-1] = r0[ALIGN16(w)-1]); luma lanes 3, 4 are x = 0, 1 and rows k-3..k-1 are y-1..y+1
                const int a3 = __builtin_amdgcn_readlane(yB, 3), a4 = __builtin_amdgcn_readlane(yB, 4);
                const int b4 = __builtin_amdgcn_readlane(yC, 4);
                const int c3 = __builtin_amdgcn_readlane(yD, 3), c4 = __builtin_amdgcn_readlane(yD, 4);
                if (x == 0) {
                    const int t0_1 = a4 + c4 + 2 * b4;
                    const int t1_0 = c3 - a3, t1_1 = c4 - a4;
                    const int t1_m1 = (w % 16 == 0) ? ((int)yflat(F, P, y - 1, w - 1) + (int)yflat(F, P, y + 1, w - 1) +
                                                       2 * (int)yflat(F, P, y, w - 1)) : 0;
                    Mc = canny_pack_f(t0_1, -(t1_m1 + 2 * t1_0 + t1_1));
                }
            }
        }
        packedPrev = canny_pack_f(gx, -gy);

        // ---- NMS of tile row r = k-6 (image row ty0+r): magnitude rows r..r+2 = MO, MN, Mc and score
        //      rows r..r+2 (maxima HO, raw Vb, maxima Hc), columns lane..lane+2, centre lane+1
        if (k >= 6) {
            const int r = k - 6, y = ty0 + r;
            const bool valid = lane < TW && (INNER || (y >= 1 && y <= h - 2 && ox >= 1 && ox <= w - 2));
            // Canny: the larger of the two neighbours along the sector's direction, packed values compared
            // through (v | 3): mag > (n >> 2) <=> (centre | 3) > (n | 3)
            const uint32_t a1 = (uint32_t)lane_next((int)MO), b1 = (uint32_t)lane_next((int)MN);
            const uint32_t c1 = (uint32_t)lane_next((int)Mc);
            const uint32_t sec = b1 & 3u;
            const uint32_t d0 = max(MN, (uint32_t)lane_next((int)b1));     // (r+1, l), (r+1, l+2)
            const uint32_t d1 = max(Mc, (uint32_t)lane_next((int)a1));     // (r, l+2), (r+2, l)
            const uint32_t d2 = max(a1, c1);                               // (r, l+1), (r+2, l+1)
            const uint32_t d3 = max(MO, (uint32_t)lane_next((int)c1));     // (r, l), (r+2, l+2)
            const uint32_t nmax = sec == 0 ? d0 : sec == 1 ? d1 : sec == 2 ? d2 : d3;
            const int edge = ((b1 | 3u) > (nmax | 3u)) ? (int)(b1 >> 2) : 0;
            const bool strongPx = valid && edge > U, edgePx = valid && (edge > U || edge > Lo);
            const uint64_t eb = __ballot(edgePx), sb = __ballot(strongPx);
            if (lane == r) { emLo = (uint32_t)eb; emHi = (uint32_t)(eb >> 32); smLo = (uint32_t)sb; smHi = (uint32_t)(sb >> 32); }
            // Harris 3x3 NMS: >= towards the upper-left half of the window, > towards the lower-right
            // the lane shifts run with every lane active: a DPP read of a lane that a short-circuited
            // && has switched off returns 0, not that lane's value
            const int c = lane_next(Vb), cRight = lane_next(c);
            const bool isMax = valid & (c >= max(HO, Vb)) & (c > max(Hc, cRight));
            const uint64_t cb = __ballot(isMax);
            if (isMax) {
                const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(cb >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)cb, 0u));
                const uint32_t bits = negKeys ? (uint32_t)(c ^ ((c >> 31) & 0x7fffffff)) : (uint32_t)c;
                cand[nc + below] = (uint64_t)(uint16_t)ox | ((uint64_t)(uint16_t)y << 16) | ((uint64_t)bits << 32);
            }
            nc += (uint32_t)__popcll(cb);
        }
        // the new rows over the oldest ones
        yA = yk; dO = dk; sO = sk;
        xx0 = hxx; xy0 = hxy; yy0 = hyy;
        MO = Mc; HO = Hc; Vb = Vc;
    };

    uint32_t pf0 = 0, pf1 = 0;
    if (INNER) { pf0 = fetch(0); pf1 = fetch(1); }
    int y0 = 0, y1 = 0, y2 = 0, y3 = 0, dA = 0, dB = 0, sA = 0, sB = 0;
    int xx0 = 0, xx1 = 0, xy0 = 0, xy1 = 0, yy0 = 0, yy1 = 0, Ha = 0, Hb = 0;
    uint32_t Ma = 0, Mb = 0;
    static_assert(YH % 4 == 2, "four steps per loop trip, two after it");
    int k = 0;
    for (; k + 4 <= YH; k += 4) {
        step(k, pf0, pf1, y0, y1, y2, y3, dA, dB, sA, sB, xx0, xx1, xy0, xy1, yy0, yy1, Ma, Mb, Ha, Hb);
        step(k + 1, pf1, pf0, y1, y2, y3, y0, dB, dA, sB, sA, xx1, xx0, xy1, xy0, yy1, yy0, Mb, Ma, Hb, Ha);
        step(k + 2, pf0, pf1, y2, y3, y0, y1, dA, dB, sA, sB, xx0, xx1, xy0, xy1, yy0, yy1, Ma, Mb, Ha, Hb);
        step(k + 3, pf1, pf0, y3, y0, y1, y2, dB, dA, sB, sA, xx1, xx0, xy1, xy0, yy1, yy0, Mb, Ma, Hb, Ha);
    }
    step(k, pf0, pf1, y0, y1, y2, y3, dA, dB, sA, sB, xx0, xx1, xy0, xy1, yy0, yy1, Ma, Mb, Ha, Hb);
    step(k + 1, pf1, pf0, y1, y2, y3, y0, dB, dA, sB, sA, xx1, xx0, xy1, xy0, yy1, yy0, Mb, Ma, Hb, Ha);

    // ---- Harris candidates: one global atomic per tile, then the buffer copied out
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (nc) {
        uint32_t base = 0;
        if (lane == 0) base = __hip_atomic_fetch_add(gmem(F.candCount), nc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        base = (uint32_t)__builtin_amdgcn_readfirstlane((int)base);
        gptr_t<uint64_t> list = gmem((uint64_t *)F.cand);
        for (uint32_t i = (uint32_t)lane; i < nc; i += 64)
            if (base + i < P.candCap) list[base + i] = cand[i];
    }

    // ---- hysteresis inside the tile.  (a) Lanes 0..31 grow the edge pixels that a strong pixel reaches
    //      (the union of the 8-connected edge components holding a strong pixel), lanes 32..63 those a
    //      tile-border edge pixel reaches.  Edge pixels in neither are weak components inside the tile
    //      (-> 0); border-touching components without a strong pixel are undecided.
    static_assert(TH <= 32 && TW <= 63, "row masks: one lane per row in each half-wave");
    const int rr = lane & 31;
    const uint64_t emine = ((uint64_t)emHi << 32) | emLo, smine = ((uint64_t)smHi << 32) | smLo;
    const uint64_t e = __shfl(emine, rr);
    uint64_t seeds;
    if (lane < 32) seeds = __shfl(smine, rr);
    else seeds = (rr == 0 || rr == th - 1) ? e : (e & (1ull | (1ull << (tw - 1))));
    const uint64_t grown = grow_rows(seeds, e, lane);
    const uint64_t other = __shfl(grown, lane ^ 32);
    const uint64_t rm = lane < 32 ? grown : other;                  // lane r (both halves): reached from strong
    const uint64_t bmAll = lane < 32 ? (other & ~grown) : 0ull;     // lane r < 32: undecided

    //  (b) settle: reached -> 255; undecided -> 127 (k_resolve decides them); other edge pixels -> 0
    for (int r = 0; r < th; r++) {
        const uint64_t R = readlane64(rm, r), B = readlane64(bmAll, r);
        const int y = ty0 + r;
        if (lane < TW && (INNER || (y >= 1 && y <= h - 2 && ox >= 1 && ox <= w - 2)))
            __builtin_amdgcn_raw_buffer_store_b8(((R >> lane) & 1ull) ? 255 : (((B >> lane) & 1ull) ? 127 : 0), edR, ox,
                                                 (int)((uint32_t)y * F.edgesStride), 0);
    }

    //  (c) the undecided components, one at a time: the first set bit in row-major order is the
    //      component's smallest index p = row*TP + column; the k-th component found gets the id
    //      1 + tile*TNB + k (each touches a border pixel no other one does); flood it inside the
    //      undecided mask.  Border tables: top row, bottom row, left column, right column of the tile
    //      (actual extent): 0 for reached pixels (the strong node), the component id for undecided
    //      ones, -1 otherwise.
    int32_t topv = -1, botv = -1, leftv = -1, rightv = -1;
    {
        const uint64_t R0 = readlane64(rm, 0), Rb = readlane64(rm, th - 1);
        if (lane < tw && ((R0 >> lane) & 1ull)) topv = 0;
        if (lane < tw && ((Rb >> lane) & 1ull)) botv = 0;
        if (lane < th && (rm & 1ull)) leftv = 0;
        if (lane < th && ((rm >> (tw - 1)) & 1ull)) rightv = 0;
    }
    gptr_t<uint32_t> ulist = gmem(F.ulist) + (size_t)tile * TS;
    uint32_t nU = 0, nComp = 0;
    uint64_t bm = bmAll;
    while (true) {
        const uint64_t rows = __ballot(bm != 0ull);
        if (!rows) break;
        const int r0 = __builtin_ctzll(rows);
        const int c0 = __builtin_ctzll(readlane64(bm, r0));
        const uint32_t gid = 1u + (uint32_t)tile * TNB + nComp;
        const uint64_t c = grow_rows(lane == r0 ? (1ull << c0) : 0ull, bm, lane);
        if (lane == 0) gmem(F.parent)[gid] = gid;
        const uint64_t C0 = readlane64(c, 0), Cb = readlane64(c, th - 1);
        if (lane < tw && ((C0 >> lane) & 1ull)) topv = (int32_t)gid;
        if (lane < tw && ((Cb >> lane) & 1ull)) botv = (int32_t)gid;
        if (lane < th && (c & 1ull)) leftv = (int32_t)gid;
        if (lane < th && ((c >> (tw - 1)) & 1ull)) rightv = (int32_t)gid;
        // list the component's pixels for k_resolve: p | k << 16
        uint64_t crow = __ballot(c != 0ull);
        while (crow) {
            const int r = __builtin_ctzll(crow);
            crow &= crow - 1;
            const uint64_t m = readlane64(c, r);
            if ((m >> lane) & 1ull) {
                const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
                ulist[nU + below] = (uint32_t)(r * TP + lane) | (nComp << 16);
            }
            nU += (uint32_t)__popcll(m);
        }
        bm &= ~c;
        nComp++;
    }
    gptr_t<int32_t> bt = gmem(F.border) + (long)tile * 2 * (TW + TH);
    if (lane < TW) { bt[lane] = topv; bt[TW + lane] = botv; }
    if (lane < TH) { bt[2 * TW + lane] = leftv; bt[2 * TW + TH + lane] = rightv; }
    if (lane == 0) gmem(F.ucount)[tile] = nU;
}

__global__ void __launch_bounds__(64 * FTW)
k_front(const FrameDesc *__restrict__ frames, GraphParams P)
{
    __shared__ uint64_t candLds[FTW][CAND_MAX];
    const FrameDesc F = frames[blockIdx.z];
    const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    const int tX = (int)blockIdx.x * FTW + wave, tY = (int)blockIdx.y;
    if (tX >= P.tilesX) return;
    const int tile = tY * P.tilesX + tX;
    const int tx0 = tX * TW, ty0 = tY * TH;
    const bool inner = tx0 >= 4 && tx0 + TW + 4 <= P.w && ty0 >= 4 && ty0 + TH + 4 <= P.h;
    if (inner) front_tile<true>(candLds[wave], F, P, tx0, ty0, tile);
    else front_tile<false>(candLds[wave], F, P, tx0, ty0, tile);
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
    const int tile = (int)blockIdx.x * 4 + __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    if (tile >= P.tilesX * P.tilesY) return;
    const int tX = tile % P.tilesX, tYi = tile / P.tilesX;
    const int tw = min(TW, P.w - tX * TW), th = min(TH, P.h - tYi * TH);
    gptr_t<const int32_t> tables = gmem((const int32_t *)F.border);
    gptr_t<const int32_t> bt = tables + (long)tile * 2 * (TW + TH);
    gptr_t<uint32_t> par = gmem(F.parent);
    // right neighbour: my right column (rows 0..th-1) against its left column, |dr| <= 1
    if (tX + 1 < P.tilesX) {
        gptr_t<const int32_t> nb = tables + (long)(tile + 1) * 2 * (TW + TH);
        border_unions(par, lane < th ? bt[2 * TW + TH + lane] : -1, lane < th ? nb[2 * TW + lane] : -1, lane);
    }
    // lower neighbour: my bottom row against its top row, |dc| <= 1 (the tile below has the same width)
    if (tYi + 1 < P.tilesY) {
        gptr_t<const int32_t> nb = tables + (long)(tile + P.tilesX) * 2 * (TW + TH);
        border_unions(par, lane < tw ? bt[TW + lane] : -1, lane < tw ? nb[lane] : -1, lane);
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
    const int w0 = ((int)blockIdx.x * 4 + __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6)) * RES_TILES;
    const int tiles = P.tilesX * P.tilesY;
    if (w0 >= tiles) return;
    // most tiles list nothing: the wave reads its tiles' counts at once and visits only the busy ones
    const uint32_t cnt = (lane < RES_TILES && w0 + lane < tiles) ? gmem(F.ucount)[w0 + lane] : 0u;
    uint64_t busy = __ballot(cnt != 0u);
    while (busy) {
        const int j = __builtin_ctzll(busy);
        busy &= busy - 1;
        const int tile = w0 + j;
        const uint32_t n = (uint32_t)__builtin_amdgcn_readlane((int)cnt, j);
        const int tx0 = (tile % P.tilesX) * TW, ty0 = (tile / P.tilesX) * TH;
        gptr_t<const uint32_t> list = gmem((const uint32_t *)F.ulist) + (size_t)tile * TS;
        for (uint32_t i = (uint32_t)lane; i < n; i += 64) {
            const uint32_t e = list[i];
            const uint32_t p = e & 0xffffu, comp = e >> 16;
            const uint32_t g = gfind(gmem(F.parent), 1u + (uint32_t)tile * TNB + comp);
            gmem(F.edges)[(long)(ty0 + (int)(p / TP)) * F.edgesStride + tx0 + (int)(p % TP)] = (g == 0) ? 255 : 0;
        }
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
    L.parent = o;    o = align256(o + ((size_t)L.tiles * TNB + 1) * 4);
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

namespace mvx {

// one stage for the frames D[0 .. nframes-1]
static void launch_stage(hipStream_t s, const GraphParams &P, const FrameDesc *D, uint32_t nframes, int stage)
{
    switch (stage) {
    case 0:
        hipLaunchKernelGGL(k_front, dim3((P.tilesX + FTW - 1) / FTW, P.tilesY, nframes), dim3(64 * FTW), 0, s, D, P);
        break;
    case 1:
        hipLaunchKernelGGL(k_border, dim3((P.tilesX * P.tilesY + 3) / 4, 1, nframes), dim3(256), 0, s, D, P);
        break;
    case 2:
        hipLaunchKernelGGL(k_resolve, dim3((P.tilesX * P.tilesY + 4 * RES_TILES - 1) / (4 * RES_TILES), 1, nframes), dim3(256), 0,
                           s, D, P);
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
}

}  // namespace mvx

extern "C" int mvx_VisionGraph_run_stage_range(mvx_stream_t stream, const mvx_vision_params_t *p, uint32_t firstFrame,
                                               uint32_t nframes, void *workspace, size_t workspaceSize, int stage)
{
    if (!p || !params_ok(*p) || nframes == 0 || !workspace || stage < 0 || stage >= MVX_VISION_STAGES ||
        firstFrame > UINT32_MAX - nframes)
        return MVX_ERROR_INVALID_PARAMETERS;
    const GraphLayout L = graph_layout(*p, firstFrame + nframes);
    if (workspaceSize < L.total) return MVX_ERROR_INVALID_PARAMETERS;
    launch_stage(stream_of(stream), graph_params(*p, L), (const FrameDesc *)workspace + firstFrame, nframes, stage);
    MVX_LAUNCH_CHECK();
    return MVX_SUCCESS;
}

extern "C" int mvx_VisionGraph_run_stage(mvx_stream_t stream, const mvx_vision_params_t *p, uint32_t nframes,
                                         void *workspace, size_t workspaceSize, int stage)
{
    return mvx_VisionGraph_run_stage_range(stream, p, 0, nframes, workspace, workspaceSize, stage);
}

// A lane: the streams the per-frame stages after the fused tile pass run on -- one for the two
// hysteresis stages, one for the pyramid levels 2.. and the Harris pick -- with the fork and join
// events, all at the default priority.  Each engine graph owns one (mvx_VisionGraph_lane_create at
// verification, destroyed with the graph), so concurrent graphs never share a side stream;
// mvx_VisionGraph_run uses one shared lane per device, created on first use and kept until the
// process ends, whose mutex orders the callers that share it.
struct mvx_vision_lane {
    hipStream_t stream = nullptr, hyst = nullptr;
    hipEvent_t fork = nullptr, join = nullptr, joinHyst = nullptr;
    std::mutex mu;
};

namespace mvx {

static int lane_init(mvx_vision_lane *L)
{
    if (hipStreamCreateWithFlags(&L->stream, hipStreamNonBlocking) != hipSuccess) return MVX_FAILURE;
    if (hipStreamCreateWithFlags(&L->hyst, hipStreamNonBlocking) != hipSuccess) return MVX_FAILURE;
    if (hipEventCreateWithFlags(&L->fork, hipEventDisableTiming) != hipSuccess) return MVX_FAILURE;
    if (hipEventCreateWithFlags(&L->join, hipEventDisableTiming) != hipSuccess) return MVX_FAILURE;
    if (hipEventCreateWithFlags(&L->joinHyst, hipEventDisableTiming) != hipSuccess) return MVX_FAILURE;
    return MVX_SUCCESS;
}

// frames per chunk of the pipelined run: by default all frames in one chunk.  Smaller chunks overlap a
// chunk's later stages with the next chunk's tile pass, but measured slower on MI355X at 4K (16 frames:
// 0.834 ms per batch in one chunk, 0.893 in two, 1.065 in four): the tile pass is bound by VALU issue,
// and the latency-bound union-find and pick waves resident beside it slow it by more than they hide.
static uint32_t mvx_vision_default_chunk(uint32_t nframes) { return nframes; }

static void lane_free(mvx_vision_lane *L)
{
    if (L->joinHyst) (void)hipEventDestroy(L->joinHyst);
    if (L->join) (void)hipEventDestroy(L->join);
    if (L->fork) (void)hipEventDestroy(L->fork);
    if (L->hyst) (void)hipStreamDestroy(L->hyst);
    if (L->stream) (void)hipStreamDestroy(L->stream);
    L->joinHyst = L->join = L->fork = nullptr;
    L->hyst = L->stream = nullptr;
}

static std::mutex g_sharedMu;
static mvx_vision_lane *g_shared[64];
static bool g_sharedTried[64];

static mvx_vision_lane *shared_lane()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return nullptr;
    std::lock_guard<std::mutex> lk(g_sharedMu);
    if (!g_sharedTried[dev]) {
        g_sharedTried[dev] = true;
        mvx_vision_lane *L = new mvx_vision_lane();
        if (lane_init(L) == MVX_SUCCESS) g_shared[dev] = L;
        else { lane_free(L); delete L; }
    }
    return g_shared[dev];
}

}  // namespace mvx

extern "C" int mvx_VisionGraph_lane_create(mvx_vision_lane_t **lane)
{
    if (!lane) return MVX_ERROR_INVALID_PARAMETERS;
    *lane = nullptr;
    mvx_vision_lane *L = new mvx_vision_lane();
    const int rc = lane_init(L);
    if (rc) { lane_free(L); delete L; return rc; }
    *lane = L;
    return MVX_SUCCESS;
}

extern "C" int mvx_VisionGraph_lane_destroy(mvx_vision_lane_t *lane)
{
    if (!lane) return MVX_SUCCESS;
    (void)hipStreamSynchronize(lane->stream);
    (void)hipStreamSynchronize(lane->hyst);
    lane_free(lane);
    delete lane;
    return MVX_SUCCESS;
}

// The frames go through in chunks of framesPerChunk: the fused tile pass of each chunk on the caller's
// stream, then -- once it has finished, while the caller's stream goes on with the next chunk's tile
// pass -- that chunk's hysteresis stages on one lane stream (the last chunk's on the caller's stream)
// and its pyramid levels and Harris pick on the other.  At the end the caller's stream waits for both
// lane streams, so the call orders like a single-stream sequence.  Fork and join are event waits, which a hipGraph capture of the caller's
// stream records as dependencies.  Once forked, the joins are recorded and waited for on every path,
// errors included, so a captured caller stream is never left forked.
extern "C" int mvx_VisionGraph_run_pipelined(mvx_stream_t stream, mvx_vision_lane_t *lane, const mvx_vision_params_t *p,
                                             uint32_t nframes, void *workspace, size_t workspaceSize,
                                             uint32_t framesPerChunk)
{
    if (!p || !params_ok(*p) || nframes == 0 || !workspace) return MVX_ERROR_INVALID_PARAMETERS;
    const GraphLayout L = graph_layout(*p, nframes);
    if (workspaceSize < L.total) return MVX_ERROR_INVALID_PARAMETERS;
    const GraphParams P = graph_params(*p, L);
    const FrameDesc *D = (const FrameDesc *)workspace;
    hipStream_t main = stream_of(stream);
    if (!lane) {
        for (int st = 0; st < MVX_VISION_STAGES; st++) launch_stage(main, P, D, nframes, st);
        MVX_LAUNCH_CHECK();
        return MVX_SUCCESS;
    }
    const uint32_t chunk = framesPerChunk ? std::min(framesPerChunk, nframes) : mvx_vision_default_chunk(nframes);
    std::lock_guard<std::mutex> lock(lane->mu);
    int rc = MVX_SUCCESS;
    bool forked = false;
    for (uint32_t f0 = 0; f0 < nframes && !rc; f0 += chunk) {
        const uint32_t n = std::min(chunk, nframes - f0);
        launch_stage(main, P, D + f0, n, 0);
        if (hipEventRecord(lane->fork, main) != hipSuccess || hipStreamWaitEvent(lane->hyst, lane->fork, 0) != hipSuccess ||
            hipStreamWaitEvent(lane->stream, lane->fork, 0) != hipSuccess) {
            rc = MVX_FAILURE;
            break;
        }
        forked = true;
        launch_stage(lane->stream, P, D + f0, n, 3);
        launch_stage(lane->stream, P, D + f0, n, 4);
        // the last chunk's hysteresis follows its tile pass on the caller's stream: no stream hop on
        // the critical path
        hipStream_t hs = f0 + n < nframes ? lane->hyst : main;
        launch_stage(hs, P, D + f0, n, 1);
        launch_stage(hs, P, D + f0, n, 2);
        if (hipGetLastError() != hipSuccess) rc = MVX_ERROR_LAUNCH;
    }
    if (forked) {
        const bool j1 = hipEventRecord(lane->joinHyst, lane->hyst) == hipSuccess &&
                        hipStreamWaitEvent(main, lane->joinHyst, 0) == hipSuccess;
        const bool j2 = hipEventRecord(lane->join, lane->stream) == hipSuccess &&
                        hipStreamWaitEvent(main, lane->join, 0) == hipSuccess;
        if (!rc && !(j1 && j2)) rc = MVX_FAILURE;
    }
    return rc;
}

extern "C" int mvx_VisionGraph_run_lane(mvx_stream_t stream, mvx_vision_lane_t *lane, const mvx_vision_params_t *p,
                                        uint32_t nframes, void *workspace, size_t workspaceSize)
{
    return mvx_VisionGraph_run_pipelined(stream, lane, p, nframes, workspace, workspaceSize, 0);
}

extern "C" int mvx_VisionGraph_run(mvx_stream_t stream, const mvx_vision_params_t *p, uint32_t nframes, void *workspace,
                                   size_t workspaceSize)
{
    return mvx_VisionGraph_run_lane(stream, shared_lane(), p, nframes, workspace, workspaceSize);
}
