// fast.hip -- FastCorners (FAST-9 on the radius-3 circle, with and without non-maximum suppression)
// for gfx950, writing the OpenVX keypoint array in the reference's raster order.
//
// The reference runs this on the CPU (its HIP FAST kernels are separate code); corners are found
// row by row, so the array order is raster order.  Here one workgroup per image row counts its
// corners, one workgroup turns the row counts into offsets, and one workgroup per row writes its
// keypoints at those offsets (wave ballots rank the corners inside each 256-pixel chunk).  With
// suppression the per-pixel strengths go to a byte map first, then the 3x3 test reads that map.
#include "mvx_common.hpp"

namespace mvx {

namespace {

// the 16-pixel circle, clockwise from the top (ago_haf_cpu_fast_corners.cpp:25-56)
__constant__ int8_t kCircle[16][2] = {{0, -3}, {1, -3}, {2, -2}, {3, -1}, {3, 0}, {3, 1}, {2, 2}, {1, 3},
                                      {0, 3}, {-1, 3}, {-2, 2}, {-3, 1}, {-3, 0}, {-3, -1}, {-2, -2}, {-1, -3}};

struct Circle {
    int c, b[16];
};

MVX_DEV Circle load_circle(const uint8_t *src, long n, uint32_t stride, int x, int y)
{
    Circle k;
    k.c = (int)ld_u8(src, n, (long)y * stride + x);
#pragma unroll
    for (int i = 0; i < 16; i++) k.b[i] = (int)ld_u8(src, n, (long)(y + kCircle[i][1]) * stride + x + kCircle[i][0]);
    return k;
}

// isCorner (:90-126): nine consecutive set bits on the circle
MVX_DEV bool nine_run(uint32_t m)
{
    if (!m) return false;
    m |= m << 16;
    uint32_t r = m;
#pragma unroll
    for (int k = 1; k < 9; k++) r &= m >> k;   // bit i set iff bits i..i+8 are all set
    return (r & 0xffffu) != 0;
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

MVX_DEV bool is_corner(const Circle &k, int t)
{
    return !early_exit(k, t) && (nine_run(plus_mask(k, t)) || nine_run(minus_mask(k, t)));
}

// checkForCornerAndGetStrength (:216-412): per sign with a nine-run at t, the largest threshold m in
// [t, 254] for which the corner test still passes (the reference bisects between t and 255, repeating the
// early exit at each probe); 0 for a sign without a run; the strength is the larger of the two, and a
// map value of 0 means no corner.  The test at m passes iff some nine-pixel arc has every difference
// d_i (b_i - c, or c - b_i) above m, i.e. m < R = max over arcs of min d_i, and the early exit does not
// fire, i.e. m <= E = min(max(|c-b0|, |c-b8|), max(|c-b4|, |c-b12|)).  Both conditions only get harder as
// m grows, so the bisection's answer is min(R - 1, E, 254), computed here directly.
MVX_DEV int strength(const Circle &k, int t)
{
    if (early_exit(k, t)) return 0;
    const int E = min(max(abs(k.c - k.b[0]), abs(k.c - k.b[8])), max(abs(k.c - k.b[4]), abs(k.c - k.b[12])));
    int best = 0;
#pragma unroll
    for (int sign = 0; sign < 2; sign++) {
        int d[16];
#pragma unroll
        for (int i = 0; i < 16; i++) d[i] = sign ? k.c - k.b[i] : k.b[i] - k.c;
        int R = -1024;
#pragma unroll
        for (int s0 = 0; s0 < 16; s0++) {
            int m = d[s0];
#pragma unroll
            for (int j = 1; j < 9; j++) m = min(m, d[(s0 + j) & 15]);
            R = max(R, m);
        }
        if (R > t) best = max(best, min(min(R - 1, E), 254));
    }
    return best;
}

struct FastArgs {
    uint32_t w, h, stride, capacity;
    const uint8_t *src;
    int t;
    float threshold;
    const uint8_t *map;   // suppression: strengths, w*h
};

// the corner decision at (x, y) of a row in 3..h-4: without suppression every x in [0, w); with it
// x in [3, w-4] and the 3x3 test on the strength map
template <bool SUPP> MVX_DEV bool corner_at(const FastArgs &a, int x, int y, float &s)
{
    if (!SUPP) {
        if (x >= (int)a.w) return false;
        s = a.threshold;
        return is_corner(load_circle(a.src, (long)a.h * a.stride, a.stride, x, y), a.t);
    }
    if (x < 3 || x >= (int)a.w - 3) return false;
    const uint8_t *p = a.map + (long)y * a.w + x;
    const int c = p[0];
    const long w = a.w;
    s = (float)c;
    return c && c >= p[-w - 1] && c >= p[-w] && c >= p[-w + 1] && c >= p[-1] && c > p[1] && c > p[w - 1] && c > p[w] &&
           c > p[w + 1];
}

// strength map over the whole plane (0 outside x in [3, w-4], y in [3, h-4])
__global__ void __launch_bounds__(256) k_fast_map(FastArgs a, uint8_t *__restrict__ map)
{
    const int x = (int)(blockIdx.x * 256 + threadIdx.x), y = (int)blockIdx.y;
    if (x >= (int)a.w) return;
    int s = 0;
    if (y >= 3 && y < (int)a.h - 3 && x >= 3 && x < (int)a.w - 3) s = strength(load_circle(a.src, (long)a.h * a.stride, a.stride, x, y), a.t);
    map[(long)y * a.w + x] = (uint8_t)s;
}

// rank of this thread's flag among the 256 of the block, and the block total
MVX_DEV uint32_t block_rank(bool f, uint32_t &total)
{
    __shared__ uint32_t waveSum[4];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint64_t m = __ballot(f);
    const uint32_t below = (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) waveSum[wave] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t base = 0;
    for (uint32_t k = 0; k < wave; k++) base += waveSum[k];
    total = waveSum[0] + waveSum[1] + waveSum[2] + waveSum[3];
    __syncthreads();
    return base + below;
}

template <bool SUPP> __global__ void __launch_bounds__(256) k_fast_count(FastArgs a, uint32_t *__restrict__ rowCount)
{
    const int y = 3 + (int)blockIdx.x;
    uint32_t n = 0;
    for (int x0 = 0; x0 < (int)a.w; x0 += 256) {
        float s;
        uint32_t total;
        block_rank(corner_at<SUPP>(a, x0 + (int)threadIdx.x, y, s), total);
        n += total;
    }
    if (threadIdx.x == 0) rowCount[blockIdx.x] = n;
}

// exclusive scan of the row counts (one workgroup); the array count (capped) and the corner count
// (uncapped) go to the caller's device words
__global__ void __launch_bounds__(1024)
k_fast_scan(const uint32_t *__restrict__ rowCount, uint32_t *__restrict__ rowOffset, uint32_t rows, uint32_t capacity,
            uint32_t *__restrict__ dstCount, uint64_t *__restrict__ numCorners)
{
    __shared__ uint32_t part[1024];
    const uint32_t per = (rows + 1023) / 1024, r0 = threadIdx.x * per;
    uint32_t sum = 0;
    for (uint32_t r = r0; r < r0 + per && r < rows; r++) sum += rowCount[r];
    part[threadIdx.x] = sum;
    __syncthreads();
    for (uint32_t d = 1; d < 1024; d <<= 1) {
        const uint32_t v = threadIdx.x >= d ? part[threadIdx.x - d] : 0u;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t run = part[threadIdx.x] - sum;
    for (uint32_t r = r0; r < r0 + per && r < rows; r++) {
        rowOffset[r] = run;
        run += rowCount[r];
    }
    if (threadIdx.x == 1023) {
        const uint32_t total = part[1023];
        *dstCount = min(total, capacity);
        if (numCorners) *numCorners = total;
    }
}

// vx_keypoint_t {x, y, strength, scale, orientation, tracking_status, error}
template <bool SUPP>
__global__ void __launch_bounds__(256)
k_fast_write(FastArgs a, const uint32_t *__restrict__ rowOffset, uint8_t *__restrict__ dst)
{
    const int y = 3 + (int)blockIdx.x;
    uint32_t base = rowOffset[blockIdx.x];
    for (int x0 = 0; x0 < (int)a.w; x0 += 256) {
        const int x = x0 + (int)threadIdx.x;
        float s = 0.0f;
        const bool f = corner_at<SUPP>(a, x, y, s);
        uint32_t total;
        const uint32_t r = block_rank(f, total);
        if (f && base + r < a.capacity) {
            uint32_t *kp = (uint32_t *)(dst + 28ull * (base + r));
            kp[0] = (uint32_t)x;
            kp[1] = (uint32_t)y;
            kp[2] = __float_as_uint(s);
            kp[3] = 0u;
            kp[4] = 0u;
            kp[5] = 1u;
            kp[6] = 0u;
        }
        base += total;
    }
}

struct FastWs {
    size_t rowCount, rowOffset, map, total;
};

FastWs fast_ws(uint32_t w, uint32_t h)
{
    FastWs L;
    const size_t rows = h > 6 ? h - 6 : 1;
    L.rowCount = 0;
    L.rowOffset = (rows * 4 + 255) & ~(size_t)255;
    L.map = L.rowOffset + ((rows * 4 + 255) & ~(size_t)255);
    L.total = L.map + (((size_t)w * h + 255) & ~(size_t)255);
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
    FastArgs a{w, h, srcStride, capacity, src, (int)(short)floorf(threshold), threshold, ws + L.map};
    hipStream_t st = stream_of(stream);
    if (SUPP) hipLaunchKernelGGL(k_fast_map, dim3((w + 255) / 256, h), dim3(256), 0, st, a, ws + L.map);
    hipLaunchKernelGGL(k_fast_count<SUPP>, dim3(rows), dim3(256), 0, st, a, (uint32_t *)(ws + L.rowCount));
    hipLaunchKernelGGL(k_fast_scan, dim3(1), dim3(1024), 0, st, (const uint32_t *)(ws + L.rowCount), (uint32_t *)(ws + L.rowOffset), rows,
                       capacity, dstCount, numCorners);
    hipLaunchKernelGGL(k_fast_write<SUPP>, dim3(rows), dim3(256), 0, st, a, (const uint32_t *)(ws + L.rowOffset), (uint8_t *)dstCorners);
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
