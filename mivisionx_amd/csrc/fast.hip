// fast.hip -- FastCorners (FAST-9 on the radius-3 circle, with and without non-maximum suppression)
// for gfx950, writing the OpenVX keypoint array in the reference's raster order.
//
// The reference runs this on the CPU (its HIP FAST kernels are separate code); corners are found
// row by row, so the array order is raster order.  Here one fused pass per 64x16 tile stages the
// tile's source bytes in LDS, finds the corners (and, with suppression, their strengths and the 3x3
// maxima) and writes a byte plane of results plus per-row counts; a scan of the counts gives each row
// its offset and one workgroup per row writes its keypoints there (wave ballots rank the corners).
#include "mvx_common.hpp"

namespace mvx {

namespace {

struct Circle {
    int c, b[16];
};

// isCorner (:90-126): nine consecutive set bits on the circle
MVX_DEV bool nine_run(uint32_t m)
{
    m |= m << 16;                               // the circle twice: runs that wrap become straight
    const uint32_t a = m & (m >> 1);            // bit i: bits i..i+1 all set
    const uint32_t b = a & (a >> 2);            //        bits i..i+3
    const uint32_t c = b & (b >> 4);            //        bits i..i+7
    return (c & (m >> 8) & 0xffffu) != 0;       //        bits i..i+8, for a start i on the circle
}

// pixels 1 and 9, or 5 and 13, both within t of the candidate
MVX_DEV bool early_exit(const Circle &k, int t)
{
    return (abs(k.c - k.b[0]) < t && abs(k.c - k.b[8]) < t) || (abs(k.c - k.b[4]) < t && abs(k.c - k.b[12]) < t);
}

MVX_DEV uint32_t plus_mask(const Circle &k, int t)
{
    uint32_t m = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) m |= (uint32_t)(k.b[i] > k.c + t) << i;
    return m;
}

MVX_DEV uint32_t minus_mask(const Circle &k, int t)
{
    uint32_t m = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) m |= (uint32_t)(k.b[i] < k.c - t) << i;
    return m;
}

// checkForCornerAndGetStrength (:216-412): per sign with a nine-run at t, the largest threshold m in
// [t, 254] for which the corner test still passes (the reference bisects between t and 255, repeating the
// early exit at each probe); 0 for a sign without a run; the strength is the larger of the two, and a
// map value of 0 means no corner.  The test at m passes iff some nine-pixel arc has every difference
// d_i (b_i - c, or c - b_i) above m, i.e. m < R = max over arcs of min d_i, and the early exit does not
// fire, i.e. m <= E = min(max(|c-b0|, |c-b8|), max(|c-b4|, |c-b12|)).  Both conditions only get harder as
// m grows, so the bisection's answer is min(R - 1, E, 254), computed here directly.  Below a zero
// threshold both signs can have runs with answers below zero; the reference then keeps the larger
// negative value, which the byte map stores modulo 256 (a sign without a run counts as 0).
MVX_DEV int strength(const Circle &k, int t)
{
    if (early_exit(k, t)) return 0;
    const int E = min(max(abs(k.c - k.b[0]), abs(k.c - k.b[8])), max(abs(k.c - k.b[4]), abs(k.c - k.b[12])));
    int s[2] = {0, 0};
#pragma unroll
    for (int sign = 0; sign < 2; sign++) {
        int d[16];
#pragma unroll
        for (int i = 0; i < 16; i++) d[i] = sign ? k.c - k.b[i] : k.b[i] - k.c;
        // minima of every circular run of 2, 4, 8, then 9 differences
        int m2[16], m4[16], m8[16];
#pragma unroll
        for (int i = 0; i < 16; i++) m2[i] = min(d[i], d[(i + 1) & 15]);
#pragma unroll
        for (int i = 0; i < 16; i++) m4[i] = min(m2[i], m2[(i + 2) & 15]);
#pragma unroll
        for (int i = 0; i < 16; i++) m8[i] = min(m4[i], m4[(i + 4) & 15]);
        int R = -1024;
#pragma unroll
        for (int i = 0; i < 16; i++) R = max(R, min(m8[i], d[(i + 8) & 15]));
        if (R > t) s[sign] = min(min(R - 1, E), 254);
    }
    return max(s[0], s[1]);
}

// ---------------------------------------------------------------- fused tile pass ------------------
// One workgroup per 64x16 tile of the image.  The tile's source bytes (rows y0-4..y0+19, columns
// x0-4..x0+67 at the reference's linear byte addresses, 0 outside the plane) go to LDS; then each of
// the four waves, on its own rows of candidate positions (lane = column, no barrier until step 3):
//   1. runs the early exit on pixels 1, 5, 9, 13 of every candidate; survivors are compacted into the
//      wave's LDS list (ballots);
//   2. runs the full nine-run test on its survivors; without suppression a corner sets its flag, with
//      it the corners are compacted again and take the strength closed form, stored (mod 256) in an
//      LDS strength map that covers the tile plus one pixel of border;
//   3. with suppression the 3x3 test runs on that map.
// The tile's results go to a byte plane (0 = none, else the strength or, without suppression, 1) and
// the per-row counts to rowCount[y - 3] (one atomic per tile row); a scan turns those into offsets and
// one workgroup per row writes its keypoints in raster order from the byte plane.
constexpr int FX = 64, FY = 16;                  // output tile
constexpr int SW = FX + 8, SH = FY + 8;          // staged source: 4 columns / rows of margin each side
static_assert(SW % 4 == 0, "staged rows are whole dwords");
constexpr int MW = FX + 2, MH = FY + 2;          // strength map: the tile plus one pixel of border
constexpr int WL = 5 * 64 + 2 * MH;              // a wave's list: its at most 5 rows, plus wave 0's 2 extra columns

struct FastArgs {
    uint32_t w, h, stride, capacity;
    const uint8_t *src;
    int t;
    float threshold;
};

// the 16-pixel circle, clockwise from the top (ago_haf_cpu_fast_corners.cpp:25-56), as offsets in the
// staged tile (row pitch SW): (0,-3) (1,-3) (2,-2) (3,-1) (3,0) (3,1) (2,2) (1,3) (0,3) (-1,3) (-2,2)
// (-3,1) (-3,0) (-3,-1) (-2,-2) (-1,-3)
__constant__ int16_t kCircleOff[16] = {-3 * SW + 0, -3 * SW + 1, -2 * SW + 2, -1 * SW + 3, 3, SW + 3, 2 * SW + 2, 3 * SW + 1,
                                       3 * SW + 0, 3 * SW - 1, 2 * SW - 2, SW - 3, -3, -SW - 3, -2 * SW - 2, -3 * SW - 1};

// append `pos` to the wave's LDS list when `keep` (n: the list's length, wave-uniform)
MVX_DEV void wave_push(bool keep, uint16_t pos, uint16_t *list, uint32_t &n)
{
    const uint64_t m = __ballot(keep);
    if (keep) list[n + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u))] = pos;
    n += (uint32_t)__popcll(m);
}

// order a wave's LDS writes before its lanes read one another's entries
MVX_DEV void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <bool SUPP>
__global__ void __launch_bounds__(256)
k_fast_tile(FastArgs a, uint8_t *__restrict__ marks, uint32_t *__restrict__ rowCount)
{
    __shared__ uint32_t stage32[SH * SW / 4];
    __shared__ uint32_t smap32[(MH * MW + 3) / 4];
    __shared__ uint16_t surv[4][WL], corn[4][WL];
    __shared__ uint32_t rowCnt[FY];
    uint8_t *stage = reinterpret_cast<uint8_t *>(stage32);
    uint8_t *smap = reinterpret_cast<uint8_t *>(smap32);
    const int tid = (int)threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int x0 = (int)blockIdx.x * FX, y0 = (int)blockIdx.y * FY;
    const long n = (long)a.h * a.stride;
    if ((((uintptr_t)a.src | a.stride) & 3u) == 0) {
        // dword loads: x0 - 4 is a multiple of 4, so every dword lies wholly inside or outside the plane
        for (int i = tid; i < SH * SW / 4; i += 256) {
            const int r = i / (SW / 4), c = i - r * (SW / 4);
            const long lin = (long)(y0 - 4 + r) * a.stride + (x0 - 4 + 4 * c);
            stage32[i] = (lin >= 0 && lin + 4 <= n) ? *reinterpret_cast<const uint32_t *>(a.src + lin) : 0u;
        }
    } else {
        for (int i = tid; i < SH * SW; i += 256) {
            const int r = i / SW, c = i - r * SW;
            stage[i] = (uint8_t)ld_u8(a.src, n, (long)(y0 - 4 + r) * a.stride + (x0 - 4 + c));
        }
    }
    for (int i = tid; i < (MH * MW + 3) / 4; i += 256) smap32[i] = 0u;
    if (tid < FY) rowCnt[tid] = 0;
    __syncthreads();
    // candidate (i, j): with suppression the map position (x0-1+i, y0-1+j), i < MW, j < MH, inside
    // [3, w-4] x [3, h-4]; without it the tile pixel (x0+i, y0+j), i < FX, j < FY, x < w, y in [3, h-4]
    constexpr int PH = SUPP ? MH : FY, D = SUPP ? 1 : 0;
    auto centre = [&](int i, int j) { return (j + 4 - D) * SW + (i + 4 - D); };
    auto early = [&](int i, int j) {
        const int x = x0 - D + i, y = y0 - D + j;
        if (!(y >= 3 && y < (int)a.h - 3 && (SUPP ? (x >= 3 && x < (int)a.w - 3) : x < (int)a.w))) return false;
        const int o = centre(i, j), c = stage[o];
        const int b0 = stage[o - 3 * SW], b4 = stage[o + 3], b8 = stage[o + 3 * SW], b12 = stage[o - 3];
        return !((abs(c - b0) < a.t && abs(c - b8) < a.t) || (abs(c - b4) < a.t && abs(c - b12) < a.t));
    };
    uint16_t *sl = surv[wave];
    uint32_t ns = 0;
    for (int j = wave; j < PH; j += 4) wave_push(early(lane, j), (uint16_t)((j << 8) | lane), sl, ns);
    if (SUPP && wave == 0) {
        // map columns 64 and 65 of every row
        const int j = lane >> 1, i = FX + (lane & 1);
        wave_push(lane < 2 * MH && early(i, j), (uint16_t)((j << 8) | i), sl, ns);
    }
    wave_sync();
    uint16_t *cl = corn[wave];
    uint32_t nc = 0;
    for (uint32_t q0 = 0; q0 < ns; q0 += 64) {
        const uint32_t q = q0 + (uint32_t)lane;
        bool corner = false;
        uint16_t pos = 0;
        if (q < ns) {
            pos = sl[q];
            const int j = pos >> 8, i = pos & 0xff;
            Circle k;
            const int o = centre(i, j);
            k.c = stage[o];
#pragma unroll
            for (int e = 0; e < 16; e++) k.b[e] = stage[o + kCircleOff[e]];
            corner = nine_run(plus_mask(k, a.t)) || nine_run(minus_mask(k, a.t));
            if (!SUPP && corner) smap[(j + 1) * MW + i + 1] = 1;
        }
        if (SUPP) wave_push(corner, pos, cl, nc);
    }
    if (SUPP) {
        wave_sync();
        for (uint32_t q = (uint32_t)lane; q < nc; q += 64) {
            const uint16_t pos = cl[q];
            const int j = pos >> 8, i = pos & 0xff;
            Circle k;
            const int o = centre(i, j);
            k.c = stage[o];
#pragma unroll
            for (int e = 0; e < 16; e++) k.b[e] = stage[o + kCircleOff[e]];
            smap[j * MW + i] = (uint8_t)strength(k, a.t);
        }
    }
    __syncthreads();
    // results: 4 pixels per thread, (x0 + 4*(tid & 15) .. +3, y0 + (tid >> 4))
    const int ty = tid >> 4, tx = (tid & 15) * 4, y = y0 + ty;
    uint32_t word = 0, cnt = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int x = x0 + tx + k;
        const uint8_t *m = smap + (ty + 1) * MW + (tx + k + 1);
        uint32_t v = 0;
        if (SUPP) {
            const int c = m[0];
            if (x >= 3 && x < (int)a.w - 3 && y >= 3 && y < (int)a.h - 3 && c && c >= m[-MW - 1] && c >= m[-MW] && c >= m[-MW + 1] &&
                c >= m[-1] && c > m[1] && c > m[MW - 1] && c > m[MW] && c > m[MW + 1])
                v = (uint32_t)c;
        } else {
            v = m[0];
        }
        cnt += v ? 1u : 0u;
        word |= v << (8 * k);
    }
    if (y < (int)a.h) {
        uint8_t *row = marks + (long)y * a.w;
        if (x0 + tx + 3 < (int)a.w && (((uintptr_t)(row + x0 + tx)) & 3u) == 0) {
            *reinterpret_cast<uint32_t *>(row + x0 + tx) = word;
        } else {
            for (int k = 0; k < 4; k++)
                if (x0 + tx + k < (int)a.w) row[x0 + tx + k] = (uint8_t)(word >> (8 * k));
        }
    }
    if (cnt) atomicAdd(&rowCnt[ty], cnt);
    __syncthreads();
    if (tid < FY && y0 + tid >= 3 && y0 + tid < (int)a.h - 3 && rowCnt[tid]) atomicAdd(&rowCount[y0 + tid - 3], rowCnt[tid]);
}

__global__ void __launch_bounds__(256) k_fast_zero(uint32_t *__restrict__ p, uint32_t n)
{
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) p[i] = 0u;
}

// exclusive prefix of v over the 256 threads of the block, and the block total (v small)
MVX_DEV uint32_t block_scan(uint32_t v, uint32_t &total)
{
    __shared__ uint32_t waveSum[4];
    const int lane = (int)(threadIdx.x & 63u), wave = (int)(threadIdx.x >> 6);
    uint32_t inc = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t t = (uint32_t)__shfl_up((int)inc, d);
        if (lane >= d) inc += t;
    }
    if (lane == 63) waveSum[wave] = inc;
    __syncthreads();
    uint32_t before = 0;
    for (int k = 0; k < wave; k++) before += waveSum[k];
    total = waveSum[0] + waveSum[1] + waveSum[2] + waveSum[3];
    __syncthreads();
    return before + inc - v;
}

// vx_keypoint_t {x, y, strength, scale, orientation, tracking_status, error}; one workgroup per row
// 3..h-4 sums the corner counts of the rows above it (its first index), reads the row's marks (16
// consecutive pixels per thread) and writes its corners from there on, in raster order.  The last row's
// workgroup, whose sum covers every row, writes the array count (capped) and the corner count (uncapped)
// to the caller's device words.
template <bool SUPP>
__global__ void __launch_bounds__(256)
k_fast_write(FastArgs a, const uint8_t *__restrict__ marks, const uint32_t *__restrict__ rowCount, uint32_t rows,
             uint32_t *__restrict__ dstCount, uint64_t *__restrict__ numCorners, uint8_t *__restrict__ dst)
{
    const int y = 3 + (int)blockIdx.x;
    uint32_t above = 0;
    for (uint32_t i = threadIdx.x; i < blockIdx.x; i += 256) above += rowCount[i];
    uint32_t base;
    block_scan(above, base);
    const uint8_t *row = marks + (long)y * a.w;
    for (int x0 = 0; x0 < (int)a.w; x0 += 256 * 16) {
        const int xs = x0 + (int)threadIdx.x * 16;
        uint32_t v[16], cnt = 0;
#pragma unroll
        for (int k = 0; k < 16; k++) {
            v[k] = xs + k < (int)a.w ? (uint32_t)row[xs + k] : 0u;
            cnt += v[k] ? 1u : 0u;
        }
        uint32_t total;
        uint32_t r = base + block_scan(cnt, total);
#pragma unroll
        for (int k = 0; k < 16; k++) {
            if (v[k]) {
                if (r < a.capacity) {
                    uint32_t *kp = (uint32_t *)(dst + 28ull * r);
                    kp[0] = (uint32_t)(xs + k);
                    kp[1] = (uint32_t)y;
                    kp[2] = __float_as_uint(SUPP ? (float)v[k] : a.threshold);
                    kp[3] = 0u;
                    kp[4] = 0u;
                    kp[5] = 1u;
                    kp[6] = 0u;
                }
                r++;
            }
        }
        base += total;
    }
    if (blockIdx.x == rows - 1 && threadIdx.x == 0) {
        *dstCount = min(base, a.capacity);
        if (numCorners) *numCorners = base;
    }
}

struct FastWs {
    size_t rowCount, marks, total;
};

FastWs fast_ws(uint32_t w, uint32_t h)
{
    FastWs L;
    const size_t rows = h > 6 ? h - 6 : 1;
    L.rowCount = 0;
    L.marks = (rows * 4 + 255) & ~(size_t)255;
    L.total = L.marks + (((size_t)w * h + 255) & ~(size_t)255);
    return L;
}

template <bool SUPP>
int fast_corners(mvx_stream_t stream, uint32_t capacity, void *dstCorners, uint32_t *dstCount, uint64_t *numCorners, uint32_t w,
                 uint32_t h, const uint8_t *src, uint32_t srcStride, float threshold, void *workspace, size_t workspaceSize)
{
    if (!dstCount || !src || !workspace || (capacity && !dstCorners)) return MVX_ERROR_INVALID_PARAMETERS;
    // the reference's suppression loop runs past its row when the image is narrower than 7
    if (w < 7 || h < 7 || w > 32767 || h > 32767) return -15;
    if (srcStride < w || ((uintptr_t)dstCorners & 3u)) return MVX_ERROR_INVALID_PARAMETERS;
    const FastWs L = fast_ws(w, h);
    if (workspaceSize < L.total) return MVX_ERROR_INVALID_PARAMETERS;
    uint8_t *ws = (uint8_t *)workspace;
    const uint32_t rows = h - 6;
    FastArgs a{w, h, srcStride, capacity, src, (int)(short)floorf(threshold), threshold};
    hipStream_t st = stream_of(stream);
    uint32_t *rowCount = (uint32_t *)(ws + L.rowCount);
    hipLaunchKernelGGL(k_fast_zero, dim3(min((rows + 255) / 256, 64u)), dim3(256), 0, st, rowCount, rows);
    hipLaunchKernelGGL(k_fast_tile<SUPP>, dim3((w + FX - 1) / FX, (h + FY - 1) / FY), dim3(256), 0, st, a, ws + L.marks, rowCount);
    hipLaunchKernelGGL(k_fast_write<SUPP>, dim3(rows), dim3(256), 0, st, a, ws + L.marks, rowCount, rows, dstCount, numCorners,
                       (uint8_t *)dstCorners);
    MVX_LAUNCH_CHECK();
    return MVX_SUCCESS;
}

}  // namespace
}  // namespace mvx

using namespace mvx;

extern "C" size_t mvx_FastCorners_workspace_size(uint32_t w, uint32_t h) { return fast_ws(w, h).total; }

extern "C" int mvx_FastCorners_XY_U8_NoSupression(mvx_stream_t stream, uint32_t capacity, void *dstCorners, uint32_t *dstCount,
                                                  uint64_t *numCorners, uint32_t w, uint32_t h, const uint8_t *src,
                                                  uint32_t srcStride, float threshold, void *workspace, size_t workspaceSize)
{
    return fast_corners<false>(stream, capacity, dstCorners, dstCount, numCorners, w, h, src, srcStride, threshold, workspace,
                               workspaceSize);
}

extern "C" int mvx_FastCorners_XY_U8_Supression(mvx_stream_t stream, uint32_t capacity, void *dstCorners, uint32_t *dstCount,
                                                uint64_t *numCorners, uint32_t w, uint32_t h, const uint8_t *src,
                                                uint32_t srcStride, float threshold, void *workspace, size_t workspaceSize)
{
    return fast_corners<true>(stream, capacity, dstCorners, dstCount, numCorners, w, h, src, srcStride, threshold, workspace,
                              workspaceSize);
}
