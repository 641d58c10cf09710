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
//   k_front     one 64x32 luma tile (+3 halo) per workgroup: RGB -> Y in LDS, then level 0, the 2:1
//               scaled image, level 1, Canny gradients/NMS/threshold with tile-local hysteresis
//               (LDS union-find), Harris products/exact integer window sums/score/NMS with atomic
//               candidate append.  Nothing but graph outputs and small per-tile tables reaches HBM.
//   k_border    unions the hysteresis components that touch tile borders, across tiles.  Strong
//               components get ids below every weak one, so after atomicMin linking a set's root is
//               strong iff the set holds a strong component -- no marking pass.
//   k_resolve   only tiles left with undecided weak pixels re-read their edge tile and settle them.
//   k_half      pyramid levels 2.. (one launch per level, all frames).
//   k_pick      Harris sort and pick, one workgroup per frame (harris.hip's algorithm).
//
// Every output is bit-exact with the reference CPU path; the equivalences used are stated where they
// are used.  Limits of the fused path (the caller falls back to the per-node kernels otherwise):
// width >= 16 and a multiple of 4 (the HG3 pitch quirk is then the identity), scaled image exactly
// half size of an even-sized input.
#include "mvx_common.hpp"
#include "pick.hpp"
#include "pyramid.hpp"
#include <algorithm>
#include <cstring>
#include <vector>

namespace mvx {

constexpr int TW = 64, TH = 32;                 // luma tile
constexpr int HY = 3;                            // luma halo: Sobel 1 + Harris window 1 + NMS 1
constexpr int YW = TW + 2 * HY, YH = TH + 2 * HY; // 70 x 38
constexpr int GW = TW + 4, GH = TH + 4;          // gradients, halo 2
constexpr int MW = TW + 2, MH = TH + 2;          // Canny magnitude / Harris score, halo 1
constexpr int TS = TW * TH;                      // pixels per tile
constexpr int FT = 256;                          // threads of the front kernel

struct FrameDesc {
    const uint8_t *rgb; uint32_t rgbStride;
    uint8_t *lvl[8]; uint32_t lvlStride[8];      // pyramid levels, lvl[0] = luma
    uint8_t *scaled; uint32_t scaledStride;
    uint8_t *edges; uint32_t edgesStride;
    void *corners; uint32_t *numItems; uint64_t *count;
    // per-frame workspace
    int32_t *border;      // tiles x 2(TW+TH) component ids of tile-border pixels (-1: none)
    uint32_t *parent;     // 2 * tiles * TS union-find slots (sparsely used)
    uint8_t *tileFlag;    // tiles: 1 when the tile holds undecided weak pixels
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
    uint32_t capacity, candCap;
};

// luma of RGB pixel (r, c), HafCpu_ColorConvert_Y_RGB (ago_haf_cpu_color_convert.cpp:2913-2990)
MVX_DEV uint32_t luma(const uint8_t *p)
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
    const long i = (long)r * P.yStride + c;
    if (i < 0 || i >= (long)P.h * P.yStride) return 0;
    r = (int)(i / P.yStride);
    c = (int)(i - (long)r * P.yStride);
    if (c >= P.w) return 0;
    return luma(F.rgb + (long)r * F.rgbStride + 3L * c);
}

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

// LDS union-find over tile pixel indices
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

MVX_DEV uint32_t gfind(const uint32_t *par, uint32_t p)
{
    uint32_t l = par[p];
    while (l != p) { p = l; l = par[p]; }
    return p;
}

MVX_DEV void gunion(uint32_t *par, uint32_t a, uint32_t b)
{
    while (true) {
        a = gfind(par, a);
        b = gfind(par, b);
        if (a == b) return;
        if (a < b) { uint32_t t = a; a = b; b = t; }
        const uint32_t old = atomicMin(&par[a], b);
        if (old == a) return;
        a = old;
    }
}



__global__ void __launch_bounds__(FT)
k_front(const FrameDesc *__restrict__ frames, GraphParams P)
{
    __shared__ uint8_t sY[YH][YW];
    __shared__ int16_t sGx[GH][GW], sGy[GH][GW];
    __shared__ uint16_t sM[MH][MW];
    __shared__ float sV[MH][MW];
    __shared__ uint8_t sS[TS];
    __shared__ int sL[TS];
    __shared__ uint8_t sFS[TS], sFB[TS];
    __shared__ uint32_t sCnt, sBase, sUnres;

    const FrameDesc &F = frames[blockIdx.z];
    const int tx0 = (int)blockIdx.x * TW, ty0 = (int)blockIdx.y * TH;
    const int tile = (int)blockIdx.y * P.tilesX + (int)blockIdx.x;
    const int t = (int)threadIdx.x;
    const int w = P.w, h = P.h;
    if (t == 0) { sCnt = 0; sUnres = 0; }

    // 1. luma window (linear-address semantics of the virtual Y plane)
    for (int i = t; i < YW * YH; i += FT) {
        const int wy = i / YW, wx = i % YW;
        sY[wy][wx] = (uint8_t)yflat(F, P, ty0 - HY + wy, tx0 - HY + wx);
    }
    __syncthreads();

    // 2. Sobel gradients (linear-address reads) at (ty0-2+j, tx0-2+i)
    for (int n = t; n < GW * GH; n += FT) {
        const int j = n / GW, i = n % GW;
        const int a0 = sY[j][i], a1 = sY[j][i + 1], a2 = sY[j][i + 2];
        const int b0 = sY[j + 1][i], b2 = sY[j + 1][i + 2];
        const int c0 = sY[j + 2][i], c1 = sY[j + 2][i + 1], c2 = sY[j + 2][i + 2];
        sGx[j][i] = (int16_t)((c2 - c0) + (a2 - a0) + ((b2 - b0) << 1));
        sGy[j][i] = (int16_t)((c0 + (c1 << 1) + c2) - (a0 + (a1 << 1) + a2));
    }
    __syncthreads();

    // 3. Canny U16 magnitude|sector and Harris score at (ty0-1+j, tx0-1+i)
    for (int n = t; n < MW * MH; n += FT) {
        const int j = n / MW, i = n % MW;
        const int y = ty0 - 1 + j, x = tx0 - 1 + i;
        uint16_t m = 0;
        float v = 0.0f;
        if (y >= 1 && y <= h - 2 && x >= 0 && x < w) {
            if (x == 0) {
                // CannySobel column 0 (ago_haf_cpu_canny.cpp:83-116 with the staged rows r0[-1] = 0 and
                // r1[-1] = r0[ALIGN16(w)-1]); window column 3 is x = 0
                const int y0w = j + 2 - 1, y1w = j + 2, y2w = j + 2 + 1;   // rows y-1, y, y+1 in sY
                const int t0_1 = (int)sY[y0w][4] + (int)sY[y2w][4] + 2 * (int)sY[y1w][4];
                const int t1_0 = (int)sY[y2w][3] - (int)sY[y0w][3];
                const int t1_1 = (int)sY[y2w][4] - (int)sY[y0w][4];
                const int t1_m1 = (w % 16 == 0) ? ((int)yflat(F, P, y - 1, w - 1) + (int)yflat(F, P, y + 1, w - 1) +
                                                   2 * (int)yflat(F, P, y, w - 1)) : 0;
                m = canny_pack_f(t0_1, -(t1_m1 + 2 * t1_0 + t1_1));
            } else {
                m = canny_pack_f(sGx[j + 1][i + 1], -(int)sGy[j + 1][i + 1]);
            }
        }
        if (y >= 1 && y <= h - 2 && x >= 1 && x <= w - 2) {
            // window sums of HG3 = {gx^2, gx gy, gy^2}; HG3 rows 0 and h-1 are never written (zero).
            // The terms are integers below 2^24 in sum, so int32 sums equal the reference's float sums.
            int sxx = 0, sxy = 0, syy = 0;
#pragma unroll
            for (int dy = -1; dy <= 1; dy++) {
                const int yy = y + dy;
                if (yy < 1 || yy > h - 2) continue;
#pragma unroll
                for (int dx = -1; dx <= 1; dx++) {
                    const int gx = sGx[j + 1 + dy][i + 1 + dx], gy = sGy[j + 1 + dy][i + 1 + dx];
                    sxx += gx * gx;
                    sxy += gx * gy;
                    syy += gy * gy;
                }
            }
            const float gx2 = (float)sxx, gxy2 = (float)sxy, gy2 = (float)syy;
            const float tr = gx2 + gy2;
            const float det = (gx2 * gy2) - (gxy2 * gxy2);
            float mc = det - (P.k * tr * tr);
            mc = __fdiv_rn(mc, P.norm);
            v = (mc > P.thr) ? mc : 0.0f;
        }
        sM[j][i] = m;
        sV[j][i] = v;
    }
    __syncthreads();

    // 4. per interior pixel: Canny NMS + threshold, Harris NMS (candidates), labels
    bool isCand[TS / FT];
    int nCand = 0;
#pragma unroll
    for (int q = 0; q < TS / FT; q++) {
        const int p = t + q * FT;
        const int ly = p / TW, lx = p % TW;
        const int y = ty0 + ly, x = tx0 + lx;
        uint8_t s = 0;
        bool cand = false;
        if (y >= 1 && y <= h - 2 && x >= 1 && x <= w - 2) {
            const int j = ly + 1, i = lx + 1;
            const uint32_t mv = sM[j][i];
            const int m = (int)(mv >> 2), a = (int)(mv & 3u);
            const int dx0 = (a == 0) ? -1 : (a == 1) ? 1 : (a == 2) ? 0 : -1;
            const int dy0 = (a == 0) ? 0 : -1;
            const int n0 = (int)(sM[j + dy0][i + dx0] >> 2), n1 = (int)(sM[j - dy0][i - dx0] >> 2);
            const int edge = (m > n0 && m > n1) ? m : 0;
            s = edge > (int)P.upper ? 255 : (edge <= (int)P.lower ? 0 : 127);
            const float c = sV[j][i];
            cand = c >= sV[j - 1][i - 1] && c >= sV[j - 1][i] && c >= sV[j - 1][i + 1] && c >= sV[j][i - 1] &&
                   c > sV[j][i + 1] && c > sV[j + 1][i - 1] && c > sV[j + 1][i] && c > sV[j + 1][i + 1];
        }
        sS[p] = s;
        sL[p] = s ? p : -1;
        sFS[p] = 0;
        sFB[p] = 0;
        isCand[q] = cand;
        nCand += cand ? 1 : 0;
    }
    // candidate append: one atomic per workgroup
    uint32_t myOff = 0;
    if (nCand) myOff = atomicAdd(&sCnt, (uint32_t)nCand);
    __syncthreads();
    if (t == 0 && sCnt) sBase = atomicAdd(F.candCount, sCnt);
    __syncthreads();
    if (nCand) {
        XYS *list = (XYS *)F.cand;
        uint32_t pos = sBase + myOff;
#pragma unroll
        for (int q = 0; q < TS / FT; q++) {
            if (!isCand[q]) continue;
            const int p = t + q * FT;
            const int ly = p / TW, lx = p % TW;
            if (pos < P.candCap) {
                XYS it;
                it.x = (int16_t)(tx0 + lx);
                it.y = (int16_t)(ty0 + ly);
                it.s = sV[ly + 1][lx + 1];
                list[pos] = it;
            }
            pos++;
        }
    }

    // 5. graph outputs that only need luma: level 0, 2:1 scaled, level 1
    {
        const int tw = min(TW, w - tx0), th = min(TH, h - ty0);
        uint8_t *l0 = F.lvl[0];
        for (int n = t; n < (TW / 4) * TH; n += FT) {
            const int ly = n / (TW / 4), lx = (n % (TW / 4)) * 4;
            if (ly >= th || lx >= tw) continue;
            uint8_t *d = l0 + (long)(ty0 + ly) * F.lvlStride[0] + tx0 + lx;
            if (lx + 4 <= tw) {
                const uint32_t v = (uint32_t)sY[ly + HY][lx + HY] | ((uint32_t)sY[ly + HY][lx + HY + 1] << 8) |
                                   ((uint32_t)sY[ly + HY][lx + HY + 2] << 16) | ((uint32_t)sY[ly + HY][lx + HY + 3] << 24);
                *reinterpret_cast<uint32_t *>(d) = v;
            } else {
                for (int u = 0; u < tw - lx; u++) d[u] = sY[ly + HY][lx + HY + u];
            }
        }
        // ScaleImage 2:1: xm = 2X, ym = 2Y, xf = yf = 128 (ago_haf_cpu_geometric.cpp:2000-2045 with
        // xscale = yscale = 2, offsets 0.5): two-stage rounded averages; tail columns single rounding
        const int body = P.dw & ~7;
        for (int n = t; n < (TW / 2) * (TH / 2); n += FT) {
            const int oy = n / (TW / 2), ox = n % (TW / 2);
            const int X = tx0 / 2 + ox, Yo = ty0 / 2 + oy;
            if (X >= P.dw || Yo >= P.dh) continue;
            const int sy = 2 * oy + HY, sx = 2 * ox + HY;
            const int p00 = sY[sy][sx], p01 = sY[sy][sx + 1], p10 = sY[sy + 1][sx], p11 = sY[sy + 1][sx + 1];
            int v;
            if (X < body) {
                const int t0 = (128 * p00 + 128 * p01 + 0x80) >> 8, t1 = (128 * p10 + 128 * p11 + 0x80) >> 8;
                v = (128 * t0 + 128 * t1 + 0x80) >> 8;
            } else {
                v = (128 * 128 * (p00 + p01 + p10 + p11) + 0x8000) >> 16;
            }
            F.scaled[(long)Yo * F.scaledStride + X] = (uint8_t)v;
        }
        // ScaleGaussianHalf_5x5 level 1 (ago_haf_cpu_pyramid.cpp:94-305): centre (2X+ox, 2Y+oy)
        if (P.levels > 1) {
            const int oxp = (w & 1) ? 0 : 1, oyp = (h & 1) ? 0 : 1;
            const int l1w = P.lvlW[1], l1h = P.lvlH[1];
            for (int n = t; n < (TW / 2) * (TH / 2); n += FT) {
                const int oy = n / (TW / 2), ox = n % (TW / 2);
                const int X = tx0 / 2 + ox, Yo = ty0 / 2 + oy;
                if (X >= l1w || Yo < 1 || Yo > l1h - 2) continue;
                const int cy = 2 * oy + oyp + HY, cx = 2 * ox + oxp + HY;   // window coordinates
                uint32_t acc = 0;
                const uint32_t kk[5] = {1, 4, 6, 4, 1};
#pragma unroll
                for (int jj = 0; jj < 5; jj++) {
                    const uint8_t *r = &sY[cy - 2 + jj][cx - 2];
                    const uint32_t hr = r[0] + 4u * r[1] + 6u * r[2] + 4u * r[3] + r[4];
                    acc += kk[jj] * hr;
                }
                F.lvl[1][(long)Yo * F.lvlStride[1] + X] = (uint8_t)((acc & 0xffffu) >> 8);
            }
        }
    }

    // 6. tile-local hysteresis: union edge pixels (8-connectivity inside the tile)
    __syncthreads();
#pragma unroll
    for (int q = 0; q < TS / FT; q++) {
        const int p = t + q * FT;
        if (!sS[p]) continue;
        const int ly = p / TW, lx = p % TW;
        if (lx + 1 < TW && sS[p + 1]) lunion(sL, p, p + 1);
        if (ly + 1 < TH) {
            if (lx > 0 && sS[p + TW - 1]) lunion(sL, p, p + TW - 1);
            if (sS[p + TW]) lunion(sL, p, p + TW);
            if (lx + 1 < TW && sS[p + TW + 1]) lunion(sL, p, p + TW + 1);
        }
    }
    __syncthreads();
    // flags on roots: holds a strong pixel / touches the tile border
    const int tw = min(TW, w - tx0), th = min(TH, h - ty0);
#pragma unroll
    for (int q = 0; q < TS / FT; q++) {
        const int p = t + q * FT;
        if (!sS[p]) continue;
        const int ly = p / TW, lx = p % TW;
        const int r = lfind(sL, p);
        sL[p] = r;   // compress (roots never change after the union pass)
        if (sS[p] == 255) sFS[r] = 1;
        if (lx == 0 || ly == 0 || lx == tw - 1 || ly == th - 1) sFB[r] = 1;
    }
    __syncthreads();
    // settle pixels; border-touching components get global ids: strong ones below every weak one
    const uint32_t weakBase = (uint32_t)(P.tilesX * P.tilesY) * TS;
    uint32_t unres = 0;
#pragma unroll
    for (int q = 0; q < TS / FT; q++) {
        const int p = t + q * FT;
        const int ly = p / TW, lx = p % TW;
        const int y = ty0 + ly, x = tx0 + lx;
        const uint8_t s = sS[p];
        if (s) {
            const int r = sL[p];
            uint8_t o;
            if (sFS[r]) o = 255;
            else if (sFB[r]) { o = 127; unres = 1; }
            else o = 0;
            if (r == p && sFB[r]) {
                const uint32_t gid = (sFS[r] ? 0u : weakBase) + (uint32_t)tile * TS + (uint32_t)p;
                F.parent[gid] = gid;
            }
            F.edges[(long)y * F.edgesStride + x] = o;
        } else if (y >= 1 && y <= h - 2 && x >= 1 && x <= w - 2) {
            F.edges[(long)y * F.edgesStride + x] = 0;
        }
    }
    if (unres) sUnres = 1;
    // border tables: top row, bottom row, left column, right column of the tile (actual extent)
    int32_t *bt = F.border + (long)tile * 2 * (TW + TH);
    for (int n = t; n < 2 * (TW + TH); n += FT) {
        int lx, ly;
        if (n < TW) { lx = n; ly = 0; }
        else if (n < 2 * TW) { lx = n - TW; ly = th - 1; }
        else if (n < 2 * TW + TH) { lx = 0; ly = n - 2 * TW; }
        else { lx = tw - 1; ly = n - 2 * TW - TH; }
        int32_t g = -1;
        if (lx < tw && ly < th) {
            const int p = ly * TW + lx;
            if (sS[p]) {
                const int r = sL[p];
                g = (int32_t)((sFS[r] ? 0u : weakBase) + (uint32_t)tile * TS + (uint32_t)r);
            }
        }
        bt[n] = g;
    }
    __syncthreads();
    if (t == 0) F.tileFlag[tile] = sUnres ? 1 : 0;
}

// union the border components of each tile with those of its right, lower and diagonal neighbours
__global__ void __launch_bounds__(256)
k_border(const FrameDesc *__restrict__ frames, GraphParams P)
{
    const FrameDesc &F = frames[blockIdx.z];
    const int tile = (int)blockIdx.x;
    const int tX = tile % P.tilesX, tYi = tile / P.tilesX;
    const int tw = min(TW, P.w - tX * TW), th = min(TH, P.h - tYi * TH);
    const int32_t *bt = F.border + (long)tile * 2 * (TW + TH);
    const int t = (int)threadIdx.x;
    // right neighbour: my right column (rows 0..th-1) against its left column, |dr| <= 1
    if (tX + 1 < P.tilesX) {
        const int32_t *nb = F.border + (long)(tile + 1) * 2 * (TW + TH);
        for (int r = t; r < th; r += 256) {
            const int32_t a = bt[2 * TW + TH + r];
            if (a < 0) continue;
            for (int d = -1; d <= 1; d++) {
                const int rr = r + d;
                if (rr < 0 || rr >= th) continue;
                const int32_t b = nb[2 * TW + rr];
                if (b >= 0) gunion(F.parent, (uint32_t)a, (uint32_t)b);
            }
        }
    }
    // lower neighbour: my bottom row against its top row, |dc| <= 1
    if (tYi + 1 < P.tilesY) {
        const int32_t *nb = F.border + (long)(tile + P.tilesX) * 2 * (TW + TH);
        const int ntw = tw;   // same column of tiles
        for (int c = t; c < tw; c += 256) {
            const int32_t a = bt[TW + c];
            if (a < 0) continue;
            for (int d = -1; d <= 1; d++) {
                const int cc = c + d;
                if (cc < 0 || cc >= ntw) continue;
                const int32_t b = nb[cc];
                if (b >= 0) gunion(F.parent, (uint32_t)a, (uint32_t)b);
            }
        }
        if (t == 0) {
            // diagonal neighbours: bottom-right corner <-> lower-right tile's top-left,
            // bottom-left corner <-> lower-left tile's top-right
            if (tX + 1 < P.tilesX) {
                const int32_t a = bt[TW + tw - 1];
                const int32_t *nd = F.border + (long)(tile + P.tilesX + 1) * 2 * (TW + TH);
                if (a >= 0 && nd[0] >= 0) gunion(F.parent, (uint32_t)a, (uint32_t)nd[0]);
            }
            if (tX > 0) {
                const int32_t a = bt[TW + 0];
                const int ltw = TW;   // the tile to the left is always full width
                const int32_t *nd = F.border + (long)(tile + P.tilesX - 1) * 2 * (TW + TH);
                if (a >= 0 && nd[ltw - 1] >= 0) gunion(F.parent, (uint32_t)a, (uint32_t)nd[ltw - 1]);
            }
        }
    }
}

// tiles holding undecided weak pixels: rebuild their components (the 127 pixels alone form exactly
// the undecided components, and LDS union-find roots are component minima, hence the same ids) and
// settle each against its global root
__global__ void __launch_bounds__(FT)
k_resolve(const FrameDesc *__restrict__ frames, GraphParams P)
{
    __shared__ int sL[TS];
    __shared__ uint8_t sS[TS];
    const FrameDesc &F = frames[blockIdx.z];
    const int tile = (int)blockIdx.y * P.tilesX + (int)blockIdx.x;
    if (!F.tileFlag[tile]) return;
    const int tx0 = (int)blockIdx.x * TW, ty0 = (int)blockIdx.y * TH;
    const int tw = min(TW, P.w - tx0), th = min(TH, P.h - ty0);
    const int t = (int)threadIdx.x;
    for (int p = t; p < TS; p += FT) {
        const int ly = p / TW, lx = p % TW;
        uint8_t s = 0;
        if (lx < tw && ly < th) s = F.edges[(long)(ty0 + ly) * F.edgesStride + tx0 + lx] == 127 ? 1 : 0;
        sS[p] = s;
        sL[p] = s ? p : -1;
    }
    __syncthreads();
    for (int p = t; p < TS; p += FT) {
        if (!sS[p]) continue;
        const int ly = p / TW, lx = p % TW;
        if (lx + 1 < TW && sS[p + 1]) lunion(sL, p, p + 1);
        if (ly + 1 < TH) {
            if (lx > 0 && sS[p + TW - 1]) lunion(sL, p, p + TW - 1);
            if (sS[p + TW]) lunion(sL, p, p + TW);
            if (lx + 1 < TW && sS[p + TW + 1]) lunion(sL, p, p + TW + 1);
        }
    }
    __syncthreads();
    const uint32_t weakBase = (uint32_t)(P.tilesX * P.tilesY) * TS;
    for (int p = t; p < TS; p += FT) {
        if (!sS[p]) continue;
        const int ly = p / TW, lx = p % TW;
        const int r = lfind(sL, p);
        const uint32_t root = gfind(F.parent, weakBase + (uint32_t)tile * TS + (uint32_t)r);
        F.edges[(long)(ty0 + ly) * F.edgesStride + tx0 + lx] = (root < weakBase) ? 255 : 0;
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
    size_t descs, frameBytes, border, parent, tileFlag, cand, candCount, pickHdr, pickWs, total;
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
    L.parent = o;    o = align256(o + (size_t)2 * L.tiles * TS * 4);
    L.tileFlag = o;  o = align256(o + L.tiles);
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
        d.tileFlag = fb + L.tileFlag;
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
        hipLaunchKernelGGL(k_border, dim3(P.tilesX * P.tilesY, 1, nframes), dim3(256), 0, s, D, P);
        break;
    case 2:
        hipLaunchKernelGGL(k_resolve, dim3(P.tilesX, P.tilesY, nframes), dim3(FT), 0, s, D, P);
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

extern "C" int mvx_VisionGraph_run(mvx_stream_t stream, const mvx_vision_params_t *p, uint32_t nframes, void *workspace,
                                   size_t workspaceSize)
{
    for (int st = 0; st < MVX_VISION_STAGES; st++) {
        const int rc = mvx_VisionGraph_run_stage(stream, p, nframes, workspace, workspaceSize, st);
        if (rc) return rc;
    }
    return MVX_SUCCESS;
}
