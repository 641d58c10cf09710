// harris.hip -- HarrisCorners (gradient and block sizes 3, 5, 7) for gfx950: HG3 Sobel products -> score ->
// 3x3 non-max suppression with raster-ordered compaction -> sort-and-pick with a minimum distance.
//
// The reference sorts and picks on the CPU (HafCpu_HarrisMergeSortAndPick_XY_XYS has no HIP
// version).  Its greedy pick visits the candidates in descending int64-key order and keeps one when
// no kept point lies closer than min_distance, i.e. it computes the lexicographically-first maximal
// independent set of the "too close" graph.  Here one workgroup sorts the keys (bitonic, in LDS when
// the list fits) and decides the set in parallel rounds: a candidate is kept once every closer,
// higher-priority candidate is rejected, rejected once one of them is kept.  That reproduces the
// sequential result exactly, without copying the list back to the host.
#include "mvx_common.hpp"
#include "pick.hpp"
#include <mutex>

namespace mvx {

inline bool plane_ok_h(const void *p, uint32_t w, uint32_t stride, uint32_t bpp)
{
    return p && aligned16(p) && (stride % 16u) == 0 && stride >= ((w * bpp + 15u) & ~15u);
}

// ---------------------------------------------------------------- HG3 plane offsets -------------
// The reference's two Harris kernels step HG3 rows by P = stride / 12 pixels, but the Sobel writer
// starts row 1 at byte `stride` (ago_haf_cpu_harris.cpp:96, :166) while the score reader puts row y at
// y*P pixels (:610-612).  Offsets in floats from the plane base:
MVX_DEV long hg3_write_off(int y, int x, uint32_t stride) { const long P = stride / 12; return (long)(stride / 4) + (long)(y - 1) * 3 * P + 3L * x; }
MVX_DEV long hg3_read_off(int y, int x, uint32_t stride) { const long P = stride / 12; return (long)y * 3 * P + 3L * x; }

// HarrisSobel_HG3_U8_3x3 (ago_haf_cpu_harris.cpp:85-169): rows 1..h-2, all columns, linear reads.
__global__ void __launch_bounds__(256)
k_harris_sobel(uint32_t w, uint32_t h, float *__restrict__ dst, uint32_t dstStride, const uint8_t *__restrict__ src,
               uint32_t srcStride)
{
    const int x = (int)(blockIdx.x * 256 + threadIdx.x);
    const int y = 1 + (int)blockIdx.y;
    if (x >= (int)w || y >= (int)h - 1) return;
    const long n = (long)h * srcStride;
    const long bm = (long)(y - 1) * srcStride + x, b0 = (long)y * srcStride + x, bp = (long)(y + 1) * srcStride + x;
    const int16_t gx = (int16_t)(((int)ld_u8(src, n, bp + 1) - (int)ld_u8(src, n, bp - 1)) +
                                 ((int)ld_u8(src, n, bm + 1) - (int)ld_u8(src, n, bm - 1)) +
                                 (((int)ld_u8(src, n, b0 + 1) - (int)ld_u8(src, n, b0 - 1)) << 1));
    const int16_t gy = (int16_t)(((int)ld_u8(src, n, bp - 1) + ((int)ld_u8(src, n, bp) << 1) + (int)ld_u8(src, n, bp + 1)) -
                                 ((int)ld_u8(src, n, bm - 1) + ((int)ld_u8(src, n, bm) << 1) + (int)ld_u8(src, n, bm + 1)));
    float *d = dst + hg3_write_off(y, x, dstStride);
    d[0] = (float)gx * (float)gx;
    d[1] = (float)gx * (float)gy;
    d[2] = (float)gy * (float)gy;
}

// HarrisScore_HVC_HG3_3x3 (ago_haf_cpu_harris.cpp:596-659): window sums j-major from 0, then
// Mc = (det - (k*tr)*tr) / norm, Vc = Mc > T ? Mc : 0; rows 0, h-1 and columns 0, w-1 are 0.
__global__ void __launch_bounds__(256)
k_harris_score(uint32_t w, uint32_t h, float *__restrict__ dst, uint32_t dstStride, const float *__restrict__ hg3,
               uint32_t hg3Stride, float k, float thr, float norm)
{
    const int x = (int)(blockIdx.x * 256 + threadIdx.x);
    const int y = (int)blockIdx.y;
    if (x >= (int)w || y >= (int)h) return;
    float *d = (float *)((uint8_t *)dst + (long)y * dstStride);
    if (y == 0 || y == (int)h - 1 || x == 0 || x == (int)w - 1) { d[x] = 0.0f; return; }
    float gx2 = 0.0f, gy2 = 0.0f, gxy2 = 0.0f;
#pragma unroll
    for (int j = -1; j <= 1; j++) {
#pragma unroll
        for (int i = -1; i <= 1; i++) {
            const float *s = hg3 + hg3_read_off(y + j, x + i, hg3Stride);
            gx2 += s[0];
            gxy2 += s[1];
            gy2 += s[2];
        }
    }
    const float tr = gx2 + gy2;
    const float det = (gx2 * gy2) - (gxy2 * gxy2);
    float mc = det - (k * tr * tr);
    mc = __fdiv_rn(mc, norm);
    d[x] = (mc > thr) ? mc : 0.0f;
}

// HarrisSobel_HG3_U8_5x5 / _7x7 (ago_haf_cpu_harris.cpp:333-594): rows R..h-R-1, all columns, linear
// reads; gx = column smoothing of the row derivatives, gy = column derivative of the row smoothings.
// The writer starts at byte R*stride and stores packed rows of w pixels (no row padding).
template <int G> struct HTaps;
template <> struct HTaps<5> { static constexpr int d[5] = {-1, -2, 0, 2, 1}, s[5] = {1, 4, 6, 4, 1}; };
template <> struct HTaps<7> { static constexpr int d[7] = {-1, -4, -5, 0, 5, 4, 1}, s[7] = {1, 6, 15, 20, 15, 6, 1}; };

template <int G>
__global__ void __launch_bounds__(256)
k_harris_sobel_g(uint32_t w, uint32_t h, float *__restrict__ dst, uint32_t dstStride, const uint8_t *__restrict__ src,
                 uint32_t srcStride)
{
    constexpr int R = G / 2;
    const int x = (int)(blockIdx.x * 256 + threadIdx.x);
    const int y = R + (int)blockIdx.y;
    if (x >= (int)w || y >= (int)h - R) return;
    const long n = (long)h * srcStride;
    int gx = 0, gy = 0;
#pragma unroll
    for (int j = 0; j < G; j++) {
        const long b = (long)(y - R + j) * srcStride + x;
        int hd = 0, hs = 0;
#pragma unroll
        for (int i = 0; i < G; i++) {
            const int p = (int)ld_u8(src, n, b + i - R);
            hd += HTaps<G>::d[i] * p;
            hs += HTaps<G>::s[i] * p;
        }
        gx += HTaps<G>::s[j] * hd;
        gy += HTaps<G>::d[j] * hs;
    }
    float *d = dst + (long)R * (dstStride / 4) + 3L * ((long)(y - R) * w + x);
    d[0] = (float)gx * (float)gx;
    d[1] = (float)gx * (float)gy;
    d[2] = (float)gy * (float)gy;
}

// HarrisScore_HVC_HG3_5x5 / _7x7 (ago_haf_cpu_harris.cpp:661-802): rows R..h-R-1, window sums j-major.
// 5x5 writes columns 2..w-3; 7x7 stores the value of column x at x-1 (x in 3..w-4) because its third
// leading zero does not advance the output pointer.  Every other pixel of the plane is 0.
template <int B>
__global__ void __launch_bounds__(256)
k_harris_score_b(uint32_t w, uint32_t h, float *__restrict__ dst, uint32_t dstStride, const float *__restrict__ hg3,
                 uint32_t hg3Stride, float k, float thr, float norm)
{
    constexpr int R = B / 2;
    const int c = (int)(blockIdx.x * 256 + threadIdx.x);
    const int y = (int)blockIdx.y;
    if (c >= (int)w || y >= (int)h) return;
    float *d = (float *)((uint8_t *)dst + (long)y * dstStride);
    const int x = B == 7 ? c + 1 : c;   // the column whose window this output holds
    if (y < R || y >= (int)h - R || x < R || x >= (int)w - R) { d[c] = 0.0f; return; }
    float gx2 = 0.0f, gy2 = 0.0f, gxy2 = 0.0f;
#pragma unroll
    for (int j = -R; j <= R; j++) {
#pragma unroll
        for (int i = -R; i <= R; i++) {
            const float *s = hg3 + hg3_read_off(y + j, x + i, hg3Stride);
            gx2 += s[0];
            gxy2 += s[1];
            gy2 += s[2];
        }
    }
    const float tr = gx2 + gy2;
    const float det = (gx2 * gy2) - (gxy2 * gxy2);
    float mc = det - (k * tr * tr);
    mc = __fdiv_rn(mc, norm);
    d[c] = (mc > thr) ? mc : 0.0f;
}

// ---------------------------------------------------------------- non-max suppression -------------
// NonMaxSupp_XY_ANY_3x3 (ago_haf_cpu_harris.cpp:882-921): keep (x, y) in [1,w-2] x [1,h-2] when
// >= the three upper and the left neighbour and > the right and the three lower ones.  Output order
// is raster order, truncated at capacity: blocks own 256-pixel row segments in raster order, count,
// get an exclusive scan of the counts, then write at their offset (ballot prefix inside the block).
MVX_DEV bool nms_keep(const float *img, uint32_t stride, int x, int y)
{
    const float *p9 = (const float *)((const uint8_t *)img + (long)(y - 1) * stride);
    const float *p0 = (const float *)((const uint8_t *)img + (long)y * stride);
    const float *p1 = (const float *)((const uint8_t *)img + (long)(y + 1) * stride);
    const float c = p0[x];
    return c >= p9[x - 1] && c >= p9[x] && c >= p9[x + 1] && c >= p0[x - 1] && c > p0[x + 1] && c > p1[x - 1] &&
           c > p1[x] && c > p1[x + 1];
}

constexpr int NMS_BLOCK = 256;

__global__ void __launch_bounds__(NMS_BLOCK)
k_nms_count(uint32_t w, uint32_t h, const float *__restrict__ img, uint32_t stride, uint32_t segs,
            uint32_t *__restrict__ blockCount)
{
    const uint32_t b = blockIdx.x;
    const int y = 1 + (int)(b / segs);
    const int x = 1 + (int)((b % segs) * NMS_BLOCK + threadIdx.x);
    const bool keep = (x < (int)w - 1) && nms_keep(img, stride, x, y);
    const int n = __syncthreads_count(keep);
    if (threadIdx.x == 0) blockCount[b] = (uint32_t)n;
}

// exclusive scan of the per-block counts (one workgroup); total stored at offsets[nb]
__global__ void __launch_bounds__(1024)
k_scan_u32(const uint32_t *__restrict__ in, uint32_t *__restrict__ out, uint32_t nb)
{
    __shared__ uint32_t part[1024];
    const uint32_t t = threadIdx.x;
    const uint32_t per = (nb + 1023) / 1024;
    const uint32_t b0 = t * per, b1 = min(b0 + per, nb);
    uint32_t s = 0;
    for (uint32_t i = b0; i < b1; i++) s += in[i];
    part[t] = s;
    __syncthreads();
    for (uint32_t off = 1; off < 1024; off <<= 1) {
        const uint32_t v = (t >= off) ? part[t - off] : 0u;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    uint32_t run = part[t] - s;   // exclusive prefix of this thread's range
    for (uint32_t i = b0; i < b1; i++) { out[i] = run; run += in[i]; }
    if (t == 1023) out[nb] = part[1023];
}

__global__ void __launch_bounds__(NMS_BLOCK)
k_nms_write(uint32_t w, uint32_t h, const float *__restrict__ img, uint32_t stride, uint32_t segs,
            const uint32_t *__restrict__ offsets, uint32_t nb, XYS *__restrict__ list, uint32_t capacity,
            uint32_t *__restrict__ count)
{
    __shared__ uint32_t waveTotal[NMS_BLOCK / 64];
    const uint32_t b = blockIdx.x;
    const int y = 1 + (int)(b / segs);
    const int x = 1 + (int)((b % segs) * NMS_BLOCK + threadIdx.x);
    const bool keep = (x < (int)w - 1) && nms_keep(img, stride, x, y);
    const uint64_t ballot = __ballot(keep);
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t below = (uint32_t)__popcll(ballot & ((1ull << lane) - 1ull));
    if (lane == 0) waveTotal[wave] = (uint32_t)__popcll(ballot);
    __syncthreads();
    uint32_t pos = offsets[b] + below;
    for (uint32_t i = 0; i < wave; i++) pos += waveTotal[i];
    if (keep && pos < capacity) {
        XYS it;
        it.x = (int16_t)x;
        it.y = (int16_t)y;
        it.s = ((const float *)((const uint8_t *)img + (long)y * stride))[x];
        list[pos] = it;
    }
    if (b == 0 && threadIdx.x == 0) *count = min(offsets[nb], capacity);
}

// ---------------------------------------------------------------- sort and pick --------------------
__global__ void __launch_bounds__(PICK_THREADS)
k_harris_pick(const XYS *__restrict__ list, uint32_t *__restrict__ listCount, uint32_t maxList, float minDistance,
              uint32_t width, uint32_t height, Keypoint *__restrict__ out, uint32_t capacity,
              uint32_t *__restrict__ outNumItems, uint64_t *__restrict__ outCount, PickWork W)
{
    harris_pick_block(list, listCount, false, maxList, minDistance, width, height, out, capacity, outNumItems, outCount, W);
}

__global__ void k_fill_u32(uint32_t *p, uint32_t v, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

}  // namespace mvx

using namespace mvx;

extern "C" int mvx_HarrisSobel_HG3_U8_3x3(mvx_stream_t stream, uint32_t w, uint32_t h, float *dst, uint32_t dstStride,
                                          const uint8_t *src, uint32_t srcStride)
{
    if (w == 0 || h == 0) return MVX_SUCCESS;
    if (!plane_ok_h(dst, w, dstStride, 12) || !plane_ok_h(src, w, srcStride, 1)) return MVX_ERROR_INVALID_PARAMETERS;
    if (h < 3) return MVX_SUCCESS;
    dim3 grid((w + 255) / 256, h - 2);
    hipLaunchKernelGGL(k_harris_sobel, grid, dim3(256), 0, stream_of(stream), w, h, dst, dstStride, src, srcStride);
    MVX_LAUNCH_CHECK();
    return MVX_SUCCESS;
}

namespace {
template <int G>
int harris_sobel_g(mvx_stream_t stream, uint32_t w, uint32_t h, float *dst, uint32_t dstStride, const uint8_t *src,
                   uint32_t srcStride)
{
    if (w == 0 || h == 0) return MVX_SUCCESS;
    if (!plane_ok_h(dst, w, dstStride, 12) || !plane_ok_h(src, w, srcStride, 1)) return MVX_ERROR_INVALID_PARAMETERS;
    if ((int)h < G) return MVX_SUCCESS;
    hipLaunchKernelGGL(k_harris_sobel_g<G>, dim3((w + 255) / 256, h - 2 * (G / 2)), dim3(256), 0, stream_of(stream), w, h, dst,
                       dstStride, src, srcStride);
    MVX_LAUNCH_CHECK();
    return MVX_SUCCESS;
}

template <int B>
int harris_score_b(mvx_stream_t stream, uint32_t w, uint32_t h, float *dst, uint32_t dstStride, const float *hg3,
                   uint32_t hg3Stride, float k, float thr, float norm)
{
    if (w == 0 || h == 0) return MVX_SUCCESS;
    if (!plane_ok_h(dst, w, dstStride, 4) || !plane_ok_h(hg3, w, hg3Stride, 12)) return MVX_ERROR_INVALID_PARAMETERS;
    // the reference writes its zero border columns and rows unconditionally: narrower planes overrun
    if (w < (uint32_t)B || h < (uint32_t)B) return -15;
    hipLaunchKernelGGL(k_harris_score_b<B>, dim3((w + 255) / 256, h), dim3(256), 0, stream_of(stream), w, h, dst, dstStride, hg3,
                       hg3Stride, k, thr, norm);
    MVX_LAUNCH_CHECK();
    return MVX_SUCCESS;
}
}  // namespace

extern "C" int mvx_HarrisSobel_HG3_U8_5x5(mvx_stream_t stream, uint32_t w, uint32_t h, float *dst, uint32_t dstStride,
                                          const uint8_t *src, uint32_t srcStride)
{
    return harris_sobel_g<5>(stream, w, h, dst, dstStride, src, srcStride);
}

extern "C" int mvx_HarrisSobel_HG3_U8_7x7(mvx_stream_t stream, uint32_t w, uint32_t h, float *dst, uint32_t dstStride,
                                          const uint8_t *src, uint32_t srcStride)
{
    return harris_sobel_g<7>(stream, w, h, dst, dstStride, src, srcStride);
}

extern "C" int mvx_HarrisScore_HVC_HG3_5x5(mvx_stream_t stream, uint32_t w, uint32_t h, float *dst, uint32_t dstStride,
                                           const float *hg3, uint32_t hg3Stride, float sensitivity, float strengthThreshold,
                                           float normFactor)
{
    return harris_score_b<5>(stream, w, h, dst, dstStride, hg3, hg3Stride, sensitivity, strengthThreshold, normFactor);
}

extern "C" int mvx_HarrisScore_HVC_HG3_7x7(mvx_stream_t stream, uint32_t w, uint32_t h, float *dst, uint32_t dstStride,
                                           const float *hg3, uint32_t hg3Stride, float sensitivity, float strengthThreshold,
                                           float normFactor)
{
    return harris_score_b<7>(stream, w, h, dst, dstStride, hg3, hg3Stride, sensitivity, strengthThreshold, normFactor);
}

extern "C" int mvx_HarrisScore_HVC_HG3_3x3(mvx_stream_t stream, uint32_t w, uint32_t h, float *dst, uint32_t dstStride,
                                           const float *hg3, uint32_t hg3Stride, float sensitivity,
                                           float strengthThreshold, float normFactor)
{
    if (w == 0 || h == 0) return MVX_SUCCESS;
    if (!plane_ok_h(dst, w, dstStride, 4) || !plane_ok_h(hg3, w, hg3Stride, 12)) return MVX_ERROR_INVALID_PARAMETERS;
    dim3 grid((w + 255) / 256, h);
    hipLaunchKernelGGL(k_harris_score, grid, dim3(256), 0, stream_of(stream), w, h, dst, dstStride, hg3, hg3Stride,
                       sensitivity, strengthThreshold, normFactor);
    MVX_LAUNCH_CHECK();
    return MVX_SUCCESS;
}

extern "C" size_t mvx_NonMaxSupp_XY_ANY_3x3_workspace_size(uint32_t w, uint32_t h)
{
    if (w < 3 || h < 3) return 256;
    const size_t segs = (w - 2 + NMS_BLOCK - 1) / NMS_BLOCK;
    const size_t nb = segs * (h - 2);
    return align256(nb * 4) + align256((nb + 1) * 4);
}

extern "C" int mvx_NonMaxSupp_XY_ANY_3x3(mvx_stream_t stream, uint32_t capacity, void *dstList, uint32_t *dstCount,
                                         uint32_t w, uint32_t h, const float *src, uint32_t srcStride, void *workspace,
                                         size_t workspaceSize)
{
    if (!dstList || !dstCount || !workspace || !plane_ok_h(src, w, srcStride, 4)) return MVX_ERROR_INVALID_PARAMETERS;
    if (workspaceSize < mvx_NonMaxSupp_XY_ANY_3x3_workspace_size(w, h)) return MVX_ERROR_INVALID_PARAMETERS;
    if (w < 3 || h < 3) {
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, stream_of(stream), dstCount, 0u, (size_t)1);
        MVX_LAUNCH_CHECK();
        return MVX_SUCCESS;
    }
    const uint32_t segs = (w - 2 + NMS_BLOCK - 1) / NMS_BLOCK;
    const uint32_t nb = segs * (h - 2);
    uint32_t *counts = (uint32_t *)workspace;
    uint32_t *offsets = (uint32_t *)((uint8_t *)workspace + align256((size_t)nb * 4));
    hipStream_t s = stream_of(stream);
    hipLaunchKernelGGL(k_nms_count, dim3(nb), dim3(NMS_BLOCK), 0, s, w, h, src, srcStride, segs, counts);
    hipLaunchKernelGGL(k_scan_u32, dim3(1), dim3(1024), 0, s, counts, offsets, nb);
    hipLaunchKernelGGL(k_nms_write, dim3(nb), dim3(NMS_BLOCK), 0, s, w, h, src, srcStride, segs, offsets, nb,
                       (XYS *)dstList, capacity, dstCount);
    MVX_LAUNCH_CHECK();
    return MVX_SUCCESS;
}

// ---------------------------------------------------------------- allocation-free NonMaxSupp ------
// The per-block counts and offsets live in one file-scope device array per GPU instead of a
// caller's workspace.  Host code never passes the array's name to a kernel: the name is a host-side
// stand-in; the device address comes from hipGetSymbolAddress, once per device, on the first call
// made with that device current.
static __device__ __align__(256) uint8_t scratch[MVX_NONMAXSUPP_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void *g_scratch[MAX_DEVICES];

int scratch_ptr(void **out)
{
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES) return MVX_FAILURE;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return MVX_FAILURE;
        g_scratch[dev] = p;
    }
    *out = g_scratch[dev];
    return MVX_SUCCESS;
}
}  // namespace

extern "C" int mvx_NonMaxSupp_XY_ANY_3x3_NoAlloc(mvx_stream_t stream, uint32_t capacity, void *dstList, uint32_t *dstCount,
                                                 uint32_t w, uint32_t h, const float *src, uint32_t srcStride)
{
    if (!dstList || !dstCount || !plane_ok_h(src, w, srcStride, 4)) return MVX_ERROR_INVALID_PARAMETERS;
    const size_t need = mvx_NonMaxSupp_XY_ANY_3x3_workspace_size(w, h);
    if (need > MVX_NONMAXSUPP_NOALLOC_SCRATCH_BYTES) return MVX_ERROR_NO_MEMORY;
    void *ws = nullptr;
    const int rc = scratch_ptr(&ws);
    if (rc != MVX_SUCCESS) return rc;
    // k_nms_count writes every count and k_scan_u32 every offset the write pass reads, so nothing
    // left by an earlier call needs clearing
    return mvx_NonMaxSupp_XY_ANY_3x3(stream, capacity, dstList, dstCount, w, h, src, srcStride, ws, need);
}

extern "C" size_t mvx_HarrisMergeSortAndPick_XY_XYS_workspace_size(uint32_t maxList)
{
    return pick_layout(maxList ? maxList : 1).total;
}

extern "C" int mvx_HarrisMergeSortAndPick_workspace_init(mvx_stream_t stream, uint32_t maxList, void *workspace,
                                                         size_t workspaceSize)
{
    const PickLayout L = pick_layout(maxList ? maxList : 1);
    if (!workspace || workspaceSize < L.total) return MVX_ERROR_INVALID_PARAMETERS;
    hipLaunchKernelGGL(k_fill_u32, dim3(256), dim3(256), 0, stream_of(stream), (uint32_t *)((uint8_t *)workspace + L.hkey),
                       HASH_EMPTY, (size_t)L.hsize);
    MVX_LAUNCH_CHECK();
    return MVX_SUCCESS;
}

extern "C" int mvx_HarrisMergeSortAndPick_XY_XYS(mvx_stream_t stream, uint32_t capacity, void *dstCorners,
                                                 uint32_t *dstNumItems, uint64_t *dstCount, const void *srcList,
                                                 const uint32_t *srcCount, uint32_t maxList, float minDistance,
                                                 uint32_t width, uint32_t height, void *workspace, size_t workspaceSize)
{
    if (!dstCorners || !dstNumItems || !dstCount || !srcList || !srcCount || !workspace) return MVX_ERROR_INVALID_PARAMETERS;
    if (minDistance <= 0.0f || width == 0 || height == 0 || width > 32767 || height > 32767) return MVX_ERROR_INVALID_PARAMETERS;
    const PickLayout L = pick_layout(maxList ? maxList : 1);
    if (workspaceSize < L.total) return MVX_ERROR_INVALID_PARAMETERS;
    const PickWork W = pick_work(workspace, maxList);
    hipLaunchKernelGGL(k_harris_pick, dim3(1), dim3(PICK_THREADS), 0, stream_of(stream), (const XYS *)srcList,
                       const_cast<uint32_t *>(srcCount), maxList, minDistance, width, height, (Keypoint *)dstCorners,
                       capacity, dstNumItems, dstCount, W);
    MVX_LAUNCH_CHECK();
    return MVX_SUCCESS;
}
