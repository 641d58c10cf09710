// lk.hip -- pyramidal Lucas-Kanade feature tracking (OpticalFlowPyrLK_XY_XY) for gfx950.
//
// HafCpu_OpticalFlowPyrLK_XY_XY_Generic (ago_haf_cpu_opticalflow.cpp:153-546) runs on the CPU only
// in the reference (ago_kernel_api.cpp:20642-20646).  Here, per pyramid level, coarse to fine:
//   k_lk_scharr       the Scharr derivative image of the old level (ComputeSharr, :59-150);
//   k_lk_scharr_tail  the state the reference's row scratch keeps after the level (see below);
//   k_lk_track        one 64-lane workgroup per keypoint: patch extraction, the 2x2 structure
//                     tensor, and the Newton iterations against the new level.
// Results are bit-identical to the reference:
//   * the reference keeps a row scratch B (u16: trow0 = B, trow1 = B + ceil8(W+2)) and the
//     derivative image D (stride = level-0 width) in the node's zero-initialised local data, and
//     reads cells of them that only earlier levels or earlier calls wrote (B[-1], B[W8],
//     B[ceil8(W+2)-1], B[ceil8(W+2)+W8]; D columns >= W8 and row 0, W8 = W & ~7).  The workspace
//     here holds the same B, D and per-point positions with the same life time, so successive
//     calls reproduce the reference's successive calls;
//   * float sums follow the reference's SSE order: per-lane partial sums over 4- or 8-column
//     groups, rows in order, scalar columns separately, combined as the reference writes it; the
//     b-vector scalar part accumulates in double;
//   * x86 conversions (truncation, 0x80000000 out of int32 range) are reproduced explicitly.
// The ordered sums are done by a handful of lanes reading per-pixel products from LDS; every other
// step is spread over the 64 lanes of the workgroup.
#include "mvx_common.hpp"
#include "pick.hpp"   // Keypoint (vx_keypoint_t)

namespace mvx {

namespace {

constexpr int LK_THREADS = 64;
constexpr int LK_MAX_WIN = 15;
constexpr int LK_MAX_PIX = LK_MAX_WIN * LK_MAX_WIN;
constexpr int LK_PPL = (LK_MAX_PIX + 63) / 64;   // window pixels per lane
constexpr int W_BITS = 14;
constexpr uint32_t TERM_EPSILON = 0xF001, TERM_BOTH = 0xF002;   // VX_TERM_CRITERIA_*
constexpr uint32_t LK_EPS_ITER_CAP = 10000;                      // EPSILON-only has no bound in the reference

MVX_DEV int32_t x86_int(float v)
{
    if (!(v > -2147483904.0f && v < 2147483648.0f)) return INT32_MIN;
    return (int32_t)v;
}

MVX_DEV int32_t sat16(int32_t v) { return min(max(v, -32768), 32767); }

MVX_DEV int32_t descale(int32_t x, int n) { return (x + (1 << (n - 1))) >> n; }

}  // namespace

struct LkLevel {
    const uint8_t *oldImg, *newImg;
    uint32_t width, height, oldStride, newStride;
};

// ComputeSharr for rows 1..H-2, columns 0..W8-1 of the old level; neighbours outside [0, W8) come
// from the row scratch B exactly as the reference's loops read them.
__global__ void __launch_bounds__(256)
k_lk_scharr(LkLevel L, int16_t *__restrict__ D, uint32_t dStride, const uint16_t *__restrict__ B)
{
    const uint32_t W8 = L.width & ~7u, T1 = (L.width + 2u + 7u) & ~7u;
    const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t y = 1 + blockIdx.y;
    if (x >= W8 || y + 1 >= L.height) return;
    const uint8_t *r0 = L.oldImg + (uint64_t)(y - 1) * L.oldStride, *r1 = r0 + L.oldStride, *r2 = r1 + L.oldStride;
    auto t0 = [&](int32_t c) -> int32_t {
        if (c < 0) return (int32_t)B[-1];
        if ((uint32_t)c >= W8) return (int32_t)B[W8];
        return (r0[c] + r2[c]) * 3 + r1[c] * 10;
    };
    auto t1 = [&](int32_t c) -> int32_t {
        if (c < 0) return (int32_t)(int16_t)B[T1 - 1];
        if ((uint32_t)c >= W8) return (int32_t)(int16_t)B[T1 + W8];
        return (int32_t)r2[c] - (int32_t)r0[c];
    };
    const int32_t xi = (int32_t)x;
    const int16_t gx = (int16_t)(t0(xi + 1) - t0(xi - 1));
    const int16_t gy = (int16_t)((t1(xi - 1) + t1(xi + 1)) * 3 + t1(xi) * 10);
    int16_t *d = D + (uint64_t)y * dStride + 2ull * x;
    *reinterpret_cast<uint32_t *>(d) = (uint32_t)(uint16_t)gx | ((uint32_t)(uint16_t)gy << 16);
}

// what the reference's row scratch holds after a level: the vertical sums of its last row (H-2)
__global__ void __launch_bounds__(256)
k_lk_scharr_tail(LkLevel L, uint16_t *__restrict__ B)
{
    const uint32_t W8 = L.width & ~7u, T1 = (L.width + 2u + 7u) & ~7u;
    const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= W8 || L.height < 3) return;
    const uint32_t y = L.height - 2;
    const uint8_t *r0 = L.oldImg + (uint64_t)(y - 1) * L.oldStride, *r1 = r0 + L.oldStride, *r2 = r1 + L.oldStride;
    B[x] = (uint16_t)((r0[x] + r2[x]) * 3 + r1[x] * 10);
    B[T1 + x] = (uint16_t)((int32_t)r2[x] - (int32_t)r0[x]);
}

struct LkParams {
    uint32_t level, levels, winsz, halfWin, dStride, iterations, termination, useEstimate;
    float ptScale, epsilon;
};

__global__ void __launch_bounds__(LK_THREADS)
k_lk_track(LkLevel L, LkParams P, const int16_t *__restrict__ D, float2 *__restrict__ next, const Keypoint *__restrict__ oldKp,
           const Keypoint *__restrict__ estKp, Keypoint *__restrict__ newKp, const uint32_t *__restrict__ oldCount,
           uint32_t *__restrict__ newCount, uint32_t capacity)
{
    __shared__ int32_t sI[LK_MAX_PIX], sIx[LK_MAX_PIX], sIy[LK_MAX_PIX];
    __shared__ float sP0[LK_MAX_PIX], sP1[LK_MAX_PIX], sP2[LK_MAX_PIX];
    __shared__ float sAcc[24];   // 0..2 results, 4..15 SSE-lane partials, 16..18 scalar partials
    const uint32_t tid = threadIdx.x;
    const uint32_t pt = blockIdx.x;
    const uint32_t n = min(*oldCount, capacity);
    if (P.level == 0 && pt == 0 && tid == 0 && newCount) *newCount = n;
    if (pt >= n) return;
    const Keypoint ok = oldKp[pt];
    Keypoint *nk = &newKp[pt];
    if (!ok.tracking_status) {
        if (tid == 0) {
            nk->x = ok.x; nk->y = ok.y; nk->strength = ok.strength; nk->tracking_status = ok.tracking_status;
            nk->scale = ok.scale; nk->error = ok.error;
        }
        return;
    }
    const int winsz = (int)P.winsz, npix = winsz * winsz;
    const bool top = P.level + 1 == P.levels;
    float prevx = (float)ok.x * P.ptScale, prevy = (float)ok.y * P.ptScale;
    float nextx, nexty;
    int32_t status;
    if (top) {
        if (P.useEstimate) {
            const Keypoint ek = estKp[pt];
            nextx = (float)ek.x * P.ptScale;
            nexty = (float)ek.y * P.ptScale;
            status = ek.tracking_status;
            if (tid == 0) { nk->strength = ek.strength; nk->tracking_status = ek.tracking_status; nk->error = ek.error; }
        } else {
            nextx = prevx;
            nexty = prevy;
            status = ok.tracking_status;
            if (tid == 0) { nk->tracking_status = ok.tracking_status; nk->strength = ok.strength; }
        }
        if (tid == 0) next[pt] = make_float2(nextx, nexty);
    } else {
        const float2 q = next[pt];
        nextx = q.x * 2.0f;
        nexty = q.y * 2.0f;
        status = nk->tracking_status;
        __syncthreads();   // every lane has read next[pt] and the status before lane 0 rewrites next[pt]
        if (tid == 0) next[pt] = make_float2(nextx, nexty);
    }
    if (!status) return;
    const float hw = (float)P.halfWin;
    prevx = prevx - hw;
    prevy = prevy - hw;
    nextx = nextx - hw;
    nexty = nexty - hw;
    const int32_t ipx = x86_int(floorf(prevx)), ipy = x86_int(floorf(prevy));
    const uint32_t dWidth = L.width - 2u, dHeight = L.height - 2u;
    if ((ipx < 0) | (ipx >= (int32_t)(dWidth - P.winsz)) | (ipy < 0) | (ipy >= (int32_t)(dHeight - P.winsz))) {
        if (P.level == 0 && tid == 0) { nk->x = x86_int(nextx); nk->y = x86_int(nexty); nk->tracking_status = 0; nk->error = 0; }
        return;
    }
    float a = prevx - (float)ipx, b = prevy - (float)ipy;
    const uint32_t IStride = L.oldStride, JStride = L.newStride, dS = P.dStride;
    const bool plain = (a == 0.0f) && (b == 0.0f);
    // ---- patch of the old image and its derivatives, and their products
    {
        const int iw00 = (int)((double)(((1.f - a) * (1.f - b)) * (float)(1 << W_BITS)) + 0.5);
        const int iw01 = (int)((double)((a * (1.f - b)) * (float)(1 << W_BITS)) + 0.5);
        const int iw10 = (int)((double)(((1.f - a) * b) * (float)(1 << W_BITS)) + 0.5);
        const int iw11 = (1 << W_BITS) - iw00 - iw01 - iw10;
        const int nA = winsz >= 4 ? (winsz - 4) / 4 + 1 : 0;   // SSE column groups of the interpolated path
        for (int p = (int)tid; p < npix; p += LK_THREADS) {
            const int y = p / winsz, x = p % winsz;
            const uint8_t *src = L.oldImg + (uint64_t)(y + ipy) * IStride + (ipx + x);
            const int16_t *ds = D + (uint64_t)(y + ipy) * dS + 2ull * (ipx + x);
            int32_t iv, ix, iy;
            if (plain) {
                iv = (int32_t)src[0] << 5;
                ix = ds[0];
                iy = ds[1];
            } else {
                iv = descale(src[0] * iw00 + src[1] * iw01 + src[IStride] * iw10 + src[IStride + 1] * iw11, W_BITS - 5);
                ix = descale(ds[0] * iw00 + ds[2] * iw01 + ds[dS] * iw10 + ds[dS + 2] * iw11, W_BITS);
                iy = descale(ds[1] * iw00 + ds[3] * iw01 + ds[dS + 1] * iw10 + ds[dS + 3] * iw11, W_BITS);
                if (x < 4 * nA) {   // the SSE columns saturate to int16 (packs_epi32)
                    iv = sat16(iv);
                    ix = sat16(ix);
                    iy = sat16(iy);
                }
            }
            sI[p] = (int32_t)(int16_t)iv;
            sIx[p] = (int32_t)(int16_t)ix;
            sIy[p] = (int32_t)(int16_t)iy;
            sP0[p] = (float)(ix * ix);
            sP1[p] = (float)(ix * iy);
            sP2[p] = (float)(iy * iy);
        }
        __syncthreads();
        if (plain) {
            // A += (float)(sum of 4 integer products) per group of x < winsz-4 step 4, then scalar columns
            if (tid < 3) {
                const int32_t *U = (tid == 1) ? sIx : (tid == 2 ? sIy : sIx);
                const int32_t *V = (tid == 0) ? sIx : sIy;
                float acc = 0.0f;
                for (int y = 0; y < winsz; y++) {
                    int x = 0;
                    for (; x < winsz - 4; x += 4) {
                        int32_t s = 0;
                        for (int k = 0; k < 4; k++) s += U[y * winsz + x + k] * V[y * winsz + x + k];
                        acc += (float)s;
                    }
                    for (; x < winsz; x++) acc += (float)(U[y * winsz + x] * V[y * winsz + x]);
                }
                sAcc[tid] = acc * (1.0f / (1 << 20));
            }
        } else {
            // 12 lanes: the 4 SSE lanes of A11, A12, A22; 3 lanes: the scalar columns
            if (tid < 12) {
                const int q = (int)tid / 4, k = (int)tid % 4;
                const float *Pq = q == 0 ? sP0 : (q == 1 ? sP1 : sP2);
                float acc = 0.0f;
                for (int y = 0; y < winsz; y++)
                    for (int g = 0; g < nA; g++) acc = acc + Pq[y * winsz + 4 * g + k];
                sAcc[4 + tid] = acc;
            } else if (tid < 15) {
                const int q = (int)tid - 12;
                const float *Pq = q == 0 ? sP0 : (q == 1 ? sP1 : sP2);
                float acc = 0.0f;
                for (int y = 0; y < winsz; y++)
                    for (int x = 4 * nA; x < winsz; x++) acc += Pq[y * winsz + x];
                sAcc[16 + q] = acc;
            }
            __syncthreads();
            if (tid < 3) {
                float acc = sAcc[16 + tid];
                if (winsz >= 4) acc += ((sAcc[4 + 4 * tid] + sAcc[5 + 4 * tid]) + sAcc[6 + 4 * tid]) + sAcc[7 + 4 * tid];
                sAcc[tid] = acc * (1.0f / (1 << 20));
            }
        }
        __syncthreads();
    }
    const float A11 = sAcc[0], A12 = sAcc[1], A22 = sAcc[2];
    float Dt = A11 * A22 - A12 * A12;
    const float minEig = (A22 + A11 - sqrtf((A11 - A22) * (A11 - A22) + 4.f * A12 * A12)) / (float)(2 * winsz * winsz);
    if (minEig < 1.0e-04F || Dt < 1.0e-07F) {
        if (P.level == 0 && tid == 0) { nk->x = x86_int(nextx); nk->y = x86_int(nexty); nk->tracking_status = 0; nk->error = 0; }
        return;
    }
    Dt = 1.f / Dt;
    // this lane's window pixels tid + 64q: row and column (-1: none)
    int pixY[LK_PPL], pixX[LK_PPL];
#pragma unroll
    for (int q = 0; q < LK_PPL; q++) {
        const int p = (int)tid + q * LK_THREADS;
        pixY[q] = p < npix ? p / winsz : -1;
        pixX[q] = p < npix ? p % winsz : 0;
    }
    float prevDx = 0.f, prevDy = 0.f, ddx = 0.f, ddy = 0.f;
    const int nB = winsz >= 8 ? (winsz - 8) / 8 + 1 : 0;   // SSE column groups of 8 in the iterations
    for (uint32_t j = 0; (j < P.iterations || P.termination == TERM_EPSILON) && j < LK_EPS_ITER_CAP;) {
        const int32_t inx = x86_int(floorf(nextx)), iny = x86_int(floorf(nexty));
        if ((inx < 0) | (inx >= (int32_t)(L.width - P.winsz)) | (iny < 0) | (iny >= (int32_t)(L.height - P.winsz))) {
            if (P.level == 0 && tid == 0) { nk->tracking_status = 0; nk->error = 0; }
            break;
        }
        a = nextx - (float)inx;
        b = nexty - (float)iny;
        const int iw00 = (int)((double)(((1.f - a) * (1.f - b)) * (float)(1 << W_BITS)) + 0.5);
        const int iw01 = (int)((double)((a * (1.f - b)) * (float)(1 << W_BITS)) + 0.5);
        const int iw10 = (int)((double)(((1.f - a) * b) * (float)(1 << W_BITS)) + 0.5);
        const int iw11 = (1 << W_BITS) - iw00 - iw01 - iw10;
        // The scalar columns' terms go straight into per-lane double sums.  Each term is a float holding an
        // integer below 2^31 in magnitude (|Ix| <= 32767 times |diff| <= 40928), and there are at most
        // 15 * 7 of them, so every partial sum is an integer below 2^38: the double sums are exact in any
        // order and equal the reference's row-major loop.
        // the lane's pixels' four source bytes are all loaded before any is used
        double sx = 0.0, sy = 0.0;
        uint32_t j00[LK_PPL], j01[LK_PPL], j10[LK_PPL], j11[LK_PPL];
#pragma unroll
        for (int q = 0; q < LK_PPL; q++) {
            j00[q] = j01[q] = j10[q] = j11[q] = 0u;
            if (pixY[q] >= 0) {
                const uint8_t *J = L.newImg + (uint64_t)(pixY[q] + iny) * JStride + (inx + pixX[q]);
                j00[q] = J[0];
                j01[q] = J[1];
                j10[q] = J[JStride];
                j11[q] = J[JStride + 1];
            }
        }
#pragma unroll
        for (int q = 0; q < LK_PPL; q++) {
            if (pixY[q] < 0) continue;
            const int p = (int)tid + q * LK_THREADS, x = pixX[q];
            int32_t t = descale((int32_t)j00[q] * iw00 + (int32_t)j01[q] * iw01 + (int32_t)j10[q] * iw10 + (int32_t)j11[q] * iw11,
                                W_BITS - 5);
            int32_t diff;
            if (x < 8 * nB) diff = sat16(sat16(t) - sI[p]);   // packs_epi32 then subs_epi16
            else diff = t - sI[p];
            const float f0 = (float)(sIx[p] * diff), f1 = (float)(sIy[p] * diff);
            if (x < 8 * nB) {
                sP0[p] = f0;
                sP1[p] = f1;
            } else {
                sx += (double)f0;
                sy += (double)f1;
            }
        }
        static_assert(LK_THREADS == 64, "the scalar-column sums are one wave's");
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) {
            sx += __shfl_xor(sx, o);
            sy += __shfl_xor(sy, o);
        }
        __syncthreads();
        if (tid < 8) {
            // qb0 lanes k: x-pair (k/2) and (4 + k/2); qb1 lanes: +2; even k: Ix, odd k: Iy.  Windows of at
            // most 15 columns have at most one group of 8 (nB <= 1); the lane's terms are read first and
            // then added in the reference's order.
            static_assert(LK_MAX_WIN < 16, "one SSE group of 8 columns per row");
            const int arr = (int)tid / 4, k = (int)tid % 4;
            const float *Pq = (k & 1) ? sP1 : sP0;
            const int c0 = 2 * arr + k / 2;
            float acc = 0.0f;
            if (nB) {
                float v0[LK_MAX_WIN], v1[LK_MAX_WIN];
#pragma unroll
                for (int y = 0; y < LK_MAX_WIN; y++) {
                    v0[y] = y < winsz ? Pq[y * winsz + c0] : 0.0f;
                    v1[y] = y < winsz ? Pq[y * winsz + 4 + c0] : 0.0f;
                }
#pragma unroll
                for (int y = 0; y < LK_MAX_WIN; y++) {
                    if (y < winsz) {
                        acc = acc + v0[y];
                        acc = acc + v1[y];
                    }
                }
            }
            sAcc[4 + tid] = acc;
        }
        __syncthreads();
        double ib1 = sx, ib2 = sy;
        if (winsz >= 8) {
            const float bb0 = sAcc[4] + sAcc[8], bb1 = sAcc[5] + sAcc[9], bb2 = sAcc[6] + sAcc[10], bb3 = sAcc[7] + sAcc[11];
            ib1 += (double)(bb0 + bb2);
            ib2 += (double)(bb1 + bb3);
        }
        const float b1 = (float)(ib1 * (double)(1.0f / (1 << 20))), b2 = (float)(ib2 * (double)(1.0f / (1 << 20)));
        const float dx = (A12 * b2 - A22 * b1) * Dt;
        const float dy = (A12 * b1 - A11 * b2) * Dt;
        nextx += dx;
        nexty += dy;
        __syncthreads();   // sP0/sP1/sAcc are rewritten by the next iteration
        if ((dx * dx + dy * dy) <= P.epsilon && (P.termination == TERM_EPSILON || P.termination == TERM_BOTH)) break;
        if (j > 0 && (double)fabsf(dx + prevDx) < 0.01 && (double)fabsf(dy + prevDy) < 0.01) {
            ddx = dx * 0.5f;
            ddy = dy * 0.5f;
            break;
        }
        prevDx = dx;
        prevDy = dy;
        j++;
    }
    if (tid == 0) {
        if (P.level == 0) {
            nk->x = x86_int(nextx + hw - ddx + 0.5f);
            nk->y = x86_int(nexty + hw - ddy + 0.5f);
        } else {
            next[pt] = make_float2(nextx + hw - ddx, nexty + hw - ddy);
        }
    }
}

struct LkLayout {
    size_t B, D, next, total;
};

inline LkLayout lk_layout(uint32_t w0, uint32_t h0, uint32_t capacity)
{
    LkLayout l;
    l.B = 0;
    l.D = align256(((size_t)2 * w0 + 64) * 2);
    l.next = align256(l.D + (size_t)w0 * h0 * 4);
    l.total = align256(l.next + (size_t)capacity * 8);
    return l;
}

}  // namespace mvx

using namespace mvx;

extern "C" size_t mvx_OpticalFlowPyrLK_workspace_size(uint32_t width0, uint32_t height0, uint32_t capacity)
{
    return lk_layout(width0, height0, capacity ? capacity : 1).total;
}

extern "C" int mvx_OpticalFlowPyrLK_XY_XY(mvx_stream_t stream, void *pNewKeypoints, uint32_t *pNewCount,
                                          const mvx_image_level_t *oldPyramid, const mvx_image_level_t *newPyramid, uint32_t levels,
                                          float pyramidScale, const void *pOldKeypoints, const uint32_t *pOldCount,
                                          const void *pEstimates, uint32_t capacity, int32_t termination, float epsilon,
                                          uint32_t numIterations, int32_t useInitialEstimate, uint32_t windowDimension,
                                          void *workspace, size_t workspaceSize)
{
    if (!pNewKeypoints || !oldPyramid || !newPyramid || !pOldKeypoints || !pOldCount || !workspace || !capacity) return MVX_ERROR_INVALID_PARAMETERS;
    if (levels < 1 || levels > 16 || windowDimension < 3 || windowDimension > (uint32_t)LK_MAX_WIN) return MVX_ERROR_INVALID_PARAMETERS;
    if (useInitialEstimate && !pEstimates) return MVX_ERROR_INVALID_PARAMETERS;
    if ((uint32_t)termination != 0xF000u && (uint32_t)termination != TERM_EPSILON && (uint32_t)termination != TERM_BOTH)
        return MVX_ERROR_INVALID_PARAMETERS;
    const uint32_t w0 = oldPyramid[0].width, h0 = oldPyramid[0].height;
    for (uint32_t l = 0; l < levels; l++) {
        const mvx_image_level_t &o = oldPyramid[l], &nw = newPyramid[l];
        // the reference's Scharr loop needs 8 columns; the window must fit inside the level
        if (!o.ptr || !nw.ptr || o.width != nw.width || o.height != nw.height || o.width < 8 || o.height < 3 ||
            o.width > w0 || o.height > h0 || o.stride < o.width || nw.stride < nw.width)
            return MVX_ERROR_INVALID_PARAMETERS;
    }
    const LkLayout lay = lk_layout(w0, h0, capacity);
    if (workspaceSize < lay.total || ((uintptr_t)workspace & 15u)) return MVX_ERROR_INVALID_PARAMETERS;
    uint8_t *ws = (uint8_t *)workspace;
    uint16_t *B = (uint16_t *)(ws + lay.B) + 8;
    int16_t *D = (int16_t *)(ws + lay.D);
    float2 *next = (float2 *)(ws + lay.next);
    const hipStream_t st = stream_of(stream);
    for (int level = (int)levels - 1; level >= 0; level--) {
        LkLevel L{oldPyramid[level].ptr, newPyramid[level].ptr, oldPyramid[level].width, oldPyramid[level].height,
                  oldPyramid[level].stride, newPyramid[level].stride};
        const uint32_t W8 = L.width & ~7u;
        hipLaunchKernelGGL(k_lk_scharr, dim3((W8 + 255) / 256, L.height - 2), dim3(256), 0, st, L, D, 2 * w0, (const uint16_t *)B);
        hipLaunchKernelGGL(k_lk_scharr_tail, dim3((W8 + 255) / 256), dim3(256), 0, st, L, B);
        LkParams P;
        P.level = (uint32_t)level;
        P.levels = levels;
        P.winsz = windowDimension;
        P.halfWin = windowDimension >> 1;
        P.dStride = 2 * w0;
        P.iterations = numIterations;
        P.termination = (uint32_t)termination;
        P.useEstimate = useInitialEstimate ? 1u : 0u;
        P.ptScale = (float)pow((double)pyramidScale, (double)level);
        P.epsilon = epsilon;
        hipLaunchKernelGGL(k_lk_track, dim3(capacity), dim3(LK_THREADS), 0, st, L, P, (const int16_t *)D, next,
                           (const Keypoint *)pOldKeypoints, (const Keypoint *)pEstimates, (Keypoint *)pNewKeypoints, pOldCount,
                           pNewCount, capacity);
        MVX_LAUNCH_CHECK();
    }
    return MVX_SUCCESS;
}
