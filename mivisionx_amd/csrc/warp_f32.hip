// warp_f32.hip -- differentiable F32 bilinear WarpPerspective and Remap, their adjoints, a fused
// L2 photometric loss with its gradient, and the F32 half-scale Gaussian pyramid with its adjoint.
//
// The reference has no F32 warp and no adjoint (its warps are U8, ago_haf_cpu_geometric.cpp:1579
// and :249); this is the differentiable counterpart used for gradient-based image alignment.  The
// sampling convention is the reference's: a destination pixel (x, y) samples the source at
// (u, v) = ((h0 x + h1 y + h2) / z, (h3 x + h4 y + h5) / z), z = h6 x + h7 y + h8 (the OpenVX matrix
// maps destination to source), bilinearly from the four pixels around (u, v); source pixels
// outside [0, sw) x [0, sh) read 0 (constant border 0), and a destination pixel with z <= 0 or a
// non-finite sample position is 0 with zero gradient.  Zero padding keeps the warp continuous in
// (u, v) everywhere, so its adjoint is the derivative almost everywhere.
//
// Adjoints, for g = dL/d out(x, y):
//   dL/d src   scatter-add of g times the four bilinear weights (hardware F32 atomics; the order of
//              additions is not fixed, so the last bit may vary between runs);
//   dL/d H     Iu = (1-fy)(p01-p00) + fy(p11-p10), Iv = (1-fx)(p10-p00) + fx(p11-p01),
//              a = g Iu / z, b = g Iv / z, c = -(a u + b v):
//              dH = (a x, a y, a, b x, b y, b, c x, c y, c), summed over the image -- per workgroup
//              in F32, then across workgroups in F64 by a second kernel in a fixed order, so dH
//              is deterministic.
//   dL/d map   (Remap) per pixel (g Iu, g Iv).
// Images of one call share a size; `count` images lie at a fixed byte distance (batch stride).
#include "mvx_common.hpp"

namespace mvx {

constexpr int WF_BX = 64, WF_BY = 4, WF_ROWS = 4;   // block 64x4 threads, each 4 rows: 64x16 px
constexpr int WF_TILE_W = WF_BX, WF_TILE_H = WF_BY * WF_ROWS;
constexpr int WF_SLOTS = 10;                        // 9 dH entries + loss

struct F32Plane {
    const float *p;
    uint32_t w, h, stride;  // stride in bytes
    size_t batch;           // bytes between images
};

MVX_DEV const float *row_of(const F32Plane &P, uint32_t img, int y)
{
    return (const float *)((const uint8_t *)P.p + img * P.batch + (size_t)y * P.stride);
}

MVX_DEV float px(const F32Plane &P, uint32_t img, int x, int y)
{
    return (x >= 0 && y >= 0 && x < (int)P.w && y < (int)P.h) ? row_of(P, img, y)[x] : 0.f;
}

// one bilinear sample with everything the adjoint needs
struct Sample {
    bool ok;
    int x0, y0;
    float fx, fy, p00, p01, p10, p11, val;
};

// sample at (bx + du, by + dv) for integers bx, by: the split into cell and fraction is taken on
// the offsets, so the fraction keeps the precision of |du| rather than of |bx + du|
MVX_DEV Sample bilinear(const F32Plane &S, uint32_t img, int bx, int by, float du, float dv)
{
    Sample s;
    s.ok = fabsf(du) < 1.0e7f && fabsf(dv) < 1.0e7f;   // also false for NaN
    if (!s.ok) { s.val = 0.f; return s; }
    const float fu = floorf(du), fv = floorf(dv);
    s.x0 = bx + (int)fu; s.y0 = by + (int)fv;
    s.fx = du - fu; s.fy = dv - fv;
    s.p00 = px(S, img, s.x0, s.y0);     s.p01 = px(S, img, s.x0 + 1, s.y0);
    s.p10 = px(S, img, s.x0, s.y0 + 1); s.p11 = px(S, img, s.x0 + 1, s.y0 + 1);
    s.val = (1.f - s.fy) * ((1.f - s.fx) * s.p00 + s.fx * s.p01) + s.fy * ((1.f - s.fx) * s.p10 + s.fx * s.p11);
    return s;
}

MVX_DEV void scatter(float *gsrc, uint32_t gStride, uint32_t w, uint32_t h, const Sample &s, float g)
{
    const float wts[4] = {(1.f - s.fx) * (1.f - s.fy), s.fx * (1.f - s.fy), (1.f - s.fx) * s.fy, s.fx * s.fy};
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int x = s.x0 + (k & 1), y = s.y0 + (k >> 1);
        if (x >= 0 && y >= 0 && x < (int)w && y < (int)h && wts[k] != 0.f)
            atomicAdd((float *)((uint8_t *)gsrc + (size_t)y * gStride) + x, g * wts[k]);
    }
}

// The homography arrives in F64 and is used in F32 as {h0 - h8, h1, h2, h3, h4 - h8, h5, h6, h7, h8}:
// the two differences are formed in F64, so a near-identity matrix loses nothing when it is rounded
// (rounding h0 and h8 separately would shift the sample by ~1e-7 x its coordinate).
struct HomF {
    float d0, h1, h2, h3, d4, h5, h6, h7, h8;
};

MVX_DEV HomF load_hom(const double *H)
{
    HomF m;
    m.d0 = (float)(H[0] - H[8]); m.h1 = (float)H[1]; m.h2 = (float)H[2];
    m.h3 = (float)H[3]; m.d4 = (float)(H[4] - H[8]); m.h5 = (float)H[5];
    m.h6 = (float)H[6]; m.h7 = (float)H[7]; m.h8 = (float)H[8];
    return m;
}

// sample position of destination pixel (X, Y) as offsets from it: du = u - X = (h0 X + h1 Y + h2 -
// X z) / z, regrouped as ((h0 - h8) X + h1 Y + h2 - X (h6 X + h7 Y)) / z (and likewise dv), which
// for near-identity homographies is far better conditioned in F32 than u itself
MVX_DEV bool project(const HomF &H, float X, float Y, float &du, float &dv, float &z)
{
    const float q = H.h6 * X + H.h7 * Y;
    z = q + H.h8;
    if (!(z > 0.f)) return false;
    du = ((H.d0 * X + H.h1 * Y + H.h2) - X * q) / z;
    dv = ((H.h3 * X + H.d4 * Y + H.h5) - Y * q) / z;
    return true;
}

// ------------------------------------------------------------------ forward kernels --------------
__global__ void __launch_bounds__(WF_BX * WF_BY)
k_warp_f32(F32Plane S, float *dst, uint32_t dw, uint32_t dh, uint32_t dStride, size_t dBatch, const double *Hs)
{
    const uint32_t img = blockIdx.z;
    const int x = blockIdx.x * WF_BX + threadIdx.x;
    const HomF H = load_hom(Hs + 9 * img);
    for (int r = 0; r < WF_ROWS; r++) {
        const int y = blockIdx.y * WF_TILE_H + r * WF_BY + threadIdx.y;
        if (x >= (int)dw || y >= (int)dh) continue;
        float du, dv, z, out = 0.f;
        if (project(H, (float)x, (float)y, du, dv, z)) out = bilinear(S, img, x, y, du, dv).val;
        ((float *)((uint8_t *)dst + img * dBatch + (size_t)y * dStride))[x] = out;
    }
}

__global__ void __launch_bounds__(WF_BX * WF_BY)
k_remap_f32(F32Plane S, float *dst, uint32_t dw, uint32_t dh, uint32_t dStride, size_t dBatch, const float *map,
            uint32_t mStride, size_t mBatch)
{
    const uint32_t img = blockIdx.z;
    const int x = blockIdx.x * WF_BX + threadIdx.x;
    for (int r = 0; r < WF_ROWS; r++) {
        const int y = blockIdx.y * WF_TILE_H + r * WF_BY + threadIdx.y;
        if (x >= (int)dw || y >= (int)dh) continue;
        const float *m = (const float *)((const uint8_t *)map + img * mBatch + (size_t)y * mStride) + 2 * x;
        ((float *)((uint8_t *)dst + img * dBatch + (size_t)y * dStride))[x] = bilinear(S, img, 0, 0, m[0], m[1]).val;
    }
}

// ------------------------------------------------------------------ adjoint kernels --------------
// block sum of WF_SLOTS values: wave shuffles, then LDS across the 4 waves; thread 0 writes
MVX_DEV void block_sum_store(float (&acc)[WF_SLOTS], float *out)
{
    __shared__ float part[WF_BY][WF_SLOTS];
    const int tid = threadIdx.y * WF_BX + threadIdx.x;
#pragma unroll
    for (int k = 0; k < WF_SLOTS; k++) {
        float v = acc[k];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        acc[k] = v;
    }
    if ((tid & 63) == 0)
#pragma unroll
        for (int k = 0; k < WF_SLOTS; k++) part[tid >> 6][k] = acc[k];
    __syncthreads();
    if (tid < WF_SLOTS) out[tid] = ((part[0][tid] + part[1][tid]) + part[2][tid]) + part[3][tid];
}

// MODE 0: g given per pixel (gradient of an arbitrary loss); MODE 1: L2 against `T`, g = 2 weight r
template <int MODE>
__global__ void __launch_bounds__(WF_BX * WF_BY)
k_warp_f32_adjoint(F32Plane S, F32Plane G, uint32_t dw, uint32_t dh, const double *Hs, float weight, float *gsrc,
                   uint32_t gsStride, size_t gsBatch, float *partials)
{
    const uint32_t img = blockIdx.z;
    const int x = blockIdx.x * WF_BX + threadIdx.x;
    const HomF H = load_hom(Hs + 9 * img);
    float acc[WF_SLOTS];
#pragma unroll
    for (int k = 0; k < WF_SLOTS; k++) acc[k] = 0.f;
    for (int r = 0; r < WF_ROWS; r++) {
        const int y = blockIdx.y * WF_TILE_H + r * WF_BY + threadIdx.y;
        if (x >= (int)dw || y >= (int)dh) continue;
        const float X = (float)x, Y = (float)y;
        float du, dv, z;
        const bool inFront = project(H, X, Y, du, dv, z);
        Sample s;
        s.ok = false;
        s.val = 0.f;
        if (inFront) s = bilinear(S, img, x, y, du, dv);
        float g;
        if (MODE == 1) {
            const float res = s.val - row_of(G, img, y)[x];
            acc[9] += weight * res * res;
            g = 2.f * weight * res;
        } else {
            g = row_of(G, img, y)[x];
        }
        if (!s.ok || g == 0.f) continue;
        const float Iu = (1.f - s.fy) * (s.p01 - s.p00) + s.fy * (s.p11 - s.p10);
        const float Iv = (1.f - s.fx) * (s.p10 - s.p00) + s.fx * (s.p11 - s.p01);
        const float u = X + du, v = Y + dv;
        const float a = g * Iu / z, b = g * Iv / z, c = -(a * u + b * v);
        acc[0] += a * X; acc[1] += a * Y; acc[2] += a;
        acc[3] += b * X; acc[4] += b * Y; acc[5] += b;
        acc[6] += c * X; acc[7] += c * Y; acc[8] += c;
        if (gsrc) scatter((float *)((uint8_t *)gsrc + img * gsBatch), gsStride, S.w, S.h, s, g);
    }
    const uint32_t blocks = gridDim.x * gridDim.y;
    block_sum_store(acc, partials + ((size_t)img * blocks + blockIdx.y * gridDim.x + blockIdx.x) * WF_SLOTS);
}

// dH[img][0..8] (and loss[img]) = F64 sums of the per-workgroup partials, in a fixed order
__global__ void __launch_bounds__(256)
k_reduce_partials(const float *partials, uint32_t blocks, double *gradH, double *loss, int accumulate)
{
    const uint32_t img = blockIdx.x;
    __shared__ double red[256];
    for (int k = 0; k < WF_SLOTS; k++) {
        double v = 0.0;
        for (uint32_t b = threadIdx.x; b < blocks; b += 256) v += (double)partials[((size_t)img * blocks + b) * WF_SLOTS + k];
        red[threadIdx.x] = v;
        __syncthreads();
        for (int off = 128; off > 0; off >>= 1) {
            if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            double *o = (k < 9) ? (gradH ? gradH + 9 * img + k : nullptr) : (loss ? loss + img : nullptr);
            if (o) *o = accumulate ? *o + red[0] : red[0];
        }
        __syncthreads();
    }
}

__global__ void __launch_bounds__(WF_BX * WF_BY)
k_remap_f32_adjoint(F32Plane S, F32Plane G, uint32_t dw, uint32_t dh, const float *map, uint32_t mStride, size_t mBatch,
                    float *gsrc, uint32_t gsStride, size_t gsBatch, float *gmap, uint32_t gmStride, size_t gmBatch)
{
    const uint32_t img = blockIdx.z;
    const int x = blockIdx.x * WF_BX + threadIdx.x;
    for (int r = 0; r < WF_ROWS; r++) {
        const int y = blockIdx.y * WF_TILE_H + r * WF_BY + threadIdx.y;
        if (x >= (int)dw || y >= (int)dh) continue;
        const float *m = (const float *)((const uint8_t *)map + img * mBatch + (size_t)y * mStride) + 2 * x;
        const Sample s = bilinear(S, img, 0, 0, m[0], m[1]);
        const float g = row_of(G, img, y)[x];
        float gu = 0.f, gv = 0.f;
        if (s.ok) {
            gu = g * ((1.f - s.fy) * (s.p01 - s.p00) + s.fy * (s.p11 - s.p10));
            gv = g * ((1.f - s.fx) * (s.p10 - s.p00) + s.fx * (s.p11 - s.p01));
            if (gsrc && g != 0.f) scatter((float *)((uint8_t *)gsrc + img * gsBatch), gsStride, S.w, S.h, s, g);
        }
        if (gmap) {
            float *o = (float *)((uint8_t *)gmap + img * gmBatch + (size_t)y * gmStride) + 2 * x;
            o[0] = gu;
            o[1] = gv;
        }
    }
}

// ------------------------------------------------------------------ F32 pyramid --------------------
// out(x, y) = sum_{i,j} w_i w_j in(clamp(2x+i-2), clamp(2y+j-2)) / 256, w = [1 4 6 4 1]
__constant__ float kPyrW[5] = {1.f, 4.f, 6.f, 4.f, 1.f};

__global__ void __launch_bounds__(WF_BX * WF_BY)
k_pyr_half_f32(F32Plane S, float *dst, uint32_t dw, uint32_t dh, uint32_t dStride, size_t dBatch)
{
    const uint32_t img = blockIdx.z;
    const int x = blockIdx.x * WF_BX + threadIdx.x;
    for (int r = 0; r < WF_ROWS; r++) {
        const int y = blockIdx.y * WF_TILE_H + r * WF_BY + threadIdx.y;
        if (x >= (int)dw || y >= (int)dh) continue;
        float acc = 0.f;
#pragma unroll
        for (int j = 0; j < 5; j++) {
            const int sy = min(max(2 * y + j - 2, 0), (int)S.h - 1);
            const float *row = row_of(S, img, sy);
            float hs = 0.f;
#pragma unroll
            for (int i = 0; i < 5; i++) hs += kPyrW[i] * row[min(max(2 * x + i - 2, 0), (int)S.w - 1)];
            acc += kPyrW[j] * hs;
        }
        ((float *)((uint8_t *)dst + img * dBatch + (size_t)y * dStride))[x] = acc * (1.f / 256.f);
    }
}

// weight with which output column x (of n_out) reads input column X (of n_in): taps whose clamped
// position is X
MVX_DEV float tap_weight(int x, int X, int n_in)
{
    float w = 0.f;
#pragma unroll
    for (int i = 0; i < 5; i++) w += (min(max(2 * x + i - 2, 0), n_in - 1) == X) ? kPyrW[i] : 0.f;
    return w;
}

// gradSrc(X, Y) += sum over the outputs that read it (a gather: no atomics)
__global__ void __launch_bounds__(WF_BX * WF_BY)
k_pyr_half_f32_adjoint(F32Plane G, float *gsrc, uint32_t sw, uint32_t sh, uint32_t gsStride, size_t gsBatch)
{
    const uint32_t img = blockIdx.z;
    const int X = blockIdx.x * WF_BX + threadIdx.x;
    for (int r = 0; r < WF_ROWS; r++) {
        const int Y = blockIdx.y * WF_TILE_H + r * WF_BY + threadIdx.y;
        if (X >= (int)sw || Y >= (int)sh) continue;
        float acc = 0.f;
        for (int y = (Y >> 1) - 1; y <= (Y >> 1) + 1; y++) {
            if (y < 0 || y >= (int)G.h) continue;
            const float wy = tap_weight(y, Y, (int)sh);
            if (wy == 0.f) continue;
            const float *row = row_of(G, img, y);
            float hs = 0.f;
            for (int x = (X >> 1) - 1; x <= (X >> 1) + 1; x++) {
                if (x < 0 || x >= (int)G.w) continue;
                hs += tap_weight(x, X, (int)sw) * row[x];
            }
            acc += wy * hs;
        }
        float *o = (float *)((uint8_t *)gsrc + img * gsBatch + (size_t)Y * gsStride) + X;
        *o += acc * (1.f / 256.f);
    }
}

inline bool f32_plane_ok(const void *p, uint32_t w, uint32_t h, uint32_t stride, size_t batch, uint32_t count, uint32_t floats = 1)
{
    if (!p || (reinterpret_cast<uintptr_t>(p) & 3u) || w == 0 || h == 0) return false;
    if (stride < w * 4u * floats || (stride & 3u)) return false;
    if (count > 1 && (batch < (size_t)stride * h || (batch & 3u))) return false;
    return true;
}

inline dim3 f32_grid(uint32_t w, uint32_t h, uint32_t count)
{
    return dim3((w + WF_TILE_W - 1) / WF_TILE_W, (h + WF_TILE_H - 1) / WF_TILE_H, count);
}

}  // namespace mvx

using namespace mvx;

extern "C" int mvx_WarpPerspective_F32_Bilinear(mvx_stream_t stream, uint32_t count, uint32_t dw, uint32_t dh, float *dst,
                                                uint32_t dstStride, size_t dstBatchStride, uint32_t sw, uint32_t sh,
                                                const float *src, uint32_t srcStride, size_t srcBatchStride, const double *H)
{
    if (count == 0) return MVX_SUCCESS;
    if (!f32_plane_ok(dst, dw, dh, dstStride, dstBatchStride, count) || !f32_plane_ok(src, sw, sh, srcStride, srcBatchStride, count) ||
        !H || count > 65535)
        return MVX_ERROR_INVALID_PARAMETERS;
    const F32Plane S{src, sw, sh, srcStride, srcBatchStride};
    hipLaunchKernelGGL(k_warp_f32, f32_grid(dw, dh, count), dim3(WF_BX, WF_BY), 0, stream_of(stream), S, dst, dw, dh, dstStride,
                       dstBatchStride, H);
    MVX_LAUNCH_CHECK();
    return MVX_SUCCESS;
}

extern "C" size_t mvx_warp_perspective_f32_workspace_size(uint32_t count, uint32_t dw, uint32_t dh)
{
    const dim3 g = f32_grid(dw, dh, count);
    return (size_t)g.x * g.y * count * WF_SLOTS * sizeof(float);
}

static int warp_adjoint_common(mvx_stream_t stream, int mode, uint32_t count, uint32_t dw, uint32_t dh, const float *g,
                               uint32_t gStride, size_t gBatch, uint32_t sw, uint32_t sh, const float *src, uint32_t srcStride,
                               size_t srcBatch, const double *H, float weight, double *loss, double *gradH, float *gradSrc,
                               uint32_t gsStride, size_t gsBatch, int accumulate, void *ws, size_t wsSize)
{
    if (count == 0) return MVX_SUCCESS;
    if (!f32_plane_ok(g, dw, dh, gStride, gBatch, count) || !f32_plane_ok(src, sw, sh, srcStride, srcBatch, count) || !H ||
        count > 65535)
        return MVX_ERROR_INVALID_PARAMETERS;
    if (gradSrc && !f32_plane_ok(gradSrc, sw, sh, gsStride, gsBatch, count)) return MVX_ERROR_INVALID_PARAMETERS;
    if (!ws || (reinterpret_cast<uintptr_t>(ws) & 3u) || wsSize < mvx_warp_perspective_f32_workspace_size(count, dw, dh))
        return MVX_ERROR_INVALID_PARAMETERS;
    const F32Plane S{src, sw, sh, srcStride, srcBatch}, G{g, dw, dh, gStride, gBatch};
    const dim3 grid = f32_grid(dw, dh, count);
    float *part = (float *)ws;
    if (mode == 1)
        hipLaunchKernelGGL(k_warp_f32_adjoint<1>, grid, dim3(WF_BX, WF_BY), 0, stream_of(stream), S, G, dw, dh, H, weight, gradSrc,
                           gsStride, gsBatch, part);
    else
        hipLaunchKernelGGL(k_warp_f32_adjoint<0>, grid, dim3(WF_BX, WF_BY), 0, stream_of(stream), S, G, dw, dh, H, weight, gradSrc,
                           gsStride, gsBatch, part);
    MVX_LAUNCH_CHECK();
    if (gradH || loss) {
        hipLaunchKernelGGL(k_reduce_partials, dim3(count), dim3(256), 0, stream_of(stream), part, grid.x * grid.y, gradH, loss,
                           accumulate);
        MVX_LAUNCH_CHECK();
    }
    return MVX_SUCCESS;
}

extern "C" int mvx_warp_perspective_f32_adjoint(mvx_stream_t stream, uint32_t count, uint32_t dw, uint32_t dh, const float *gradOut,
                                                uint32_t gStride, size_t gBatchStride, uint32_t sw, uint32_t sh, const float *src,
                                                uint32_t srcStride, size_t srcBatchStride, const double *H, float *gradSrc,
                                                uint32_t gsStride, size_t gsBatchStride, double *gradH, void *workspace,
                                                size_t workspaceSize)
{
    return warp_adjoint_common(stream, 0, count, dw, dh, gradOut, gStride, gBatchStride, sw, sh, src, srcStride, srcBatchStride, H,
                               1.f, nullptr, gradH, gradSrc, gsStride, gsBatchStride, 0, workspace, workspaceSize);
}

extern "C" int mvx_warp_perspective_f32_l2(mvx_stream_t stream, uint32_t count, uint32_t dw, uint32_t dh, const float *target,
                                           uint32_t tStride, size_t tBatchStride, uint32_t sw, uint32_t sh, const float *src,
                                           uint32_t srcStride, size_t srcBatchStride, const double *H, float weight, double *loss,
                                           double *gradH, float *gradSrc, uint32_t gsStride, size_t gsBatchStride, int accumulate,
                                           void *workspace, size_t workspaceSize)
{
    return warp_adjoint_common(stream, 1, count, dw, dh, target, tStride, tBatchStride, sw, sh, src, srcStride, srcBatchStride, H,
                               weight, loss, gradH, gradSrc, gsStride, gsBatchStride, accumulate, workspace, workspaceSize);
}

extern "C" int mvx_Remap_F32_Bilinear(mvx_stream_t stream, uint32_t count, uint32_t dw, uint32_t dh, float *dst, uint32_t dstStride,
                                      size_t dstBatchStride, uint32_t sw, uint32_t sh, const float *src, uint32_t srcStride,
                                      size_t srcBatchStride, const float *map, uint32_t mapStride, size_t mapBatchStride)
{
    if (count == 0) return MVX_SUCCESS;
    if (!f32_plane_ok(dst, dw, dh, dstStride, dstBatchStride, count) || !f32_plane_ok(src, sw, sh, srcStride, srcBatchStride, count) ||
        !f32_plane_ok(map, dw, dh, mapStride, mapBatchStride, count, 2) || count > 65535)
        return MVX_ERROR_INVALID_PARAMETERS;
    const F32Plane S{src, sw, sh, srcStride, srcBatchStride};
    hipLaunchKernelGGL(k_remap_f32, f32_grid(dw, dh, count), dim3(WF_BX, WF_BY), 0, stream_of(stream), S, dst, dw, dh, dstStride,
                       dstBatchStride, map, mapStride, mapBatchStride);
    MVX_LAUNCH_CHECK();
    return MVX_SUCCESS;
}

extern "C" int mvx_remap_f32_adjoint(mvx_stream_t stream, uint32_t count, uint32_t dw, uint32_t dh, const float *gradOut,
                                     uint32_t gStride, size_t gBatchStride, uint32_t sw, uint32_t sh, const float *src,
                                     uint32_t srcStride, size_t srcBatchStride, const float *map, uint32_t mapStride,
                                     size_t mapBatchStride, float *gradSrc, uint32_t gsStride, size_t gsBatchStride, float *gradMap,
                                     uint32_t gmStride, size_t gmBatchStride)
{
    if (count == 0) return MVX_SUCCESS;
    if (!f32_plane_ok(gradOut, dw, dh, gStride, gBatchStride, count) || !f32_plane_ok(src, sw, sh, srcStride, srcBatchStride, count) ||
        !f32_plane_ok(map, dw, dh, mapStride, mapBatchStride, count, 2) || count > 65535)
        return MVX_ERROR_INVALID_PARAMETERS;
    if (gradSrc && !f32_plane_ok(gradSrc, sw, sh, gsStride, gsBatchStride, count)) return MVX_ERROR_INVALID_PARAMETERS;
    if (gradMap && !f32_plane_ok(gradMap, dw, dh, gmStride, gmBatchStride, count, 2)) return MVX_ERROR_INVALID_PARAMETERS;
    const F32Plane S{src, sw, sh, srcStride, srcBatchStride}, G{gradOut, dw, dh, gStride, gBatchStride};
    hipLaunchKernelGGL(k_remap_f32_adjoint, f32_grid(dw, dh, count), dim3(WF_BX, WF_BY), 0, stream_of(stream), S, G, dw, dh, map,
                       mapStride, mapBatchStride, gradSrc, gsStride, gsBatchStride, gradMap, gmStride, gmBatchStride);
    MVX_LAUNCH_CHECK();
    return MVX_SUCCESS;
}

extern "C" int mvx_PyramidHalf_F32(mvx_stream_t stream, uint32_t count, uint32_t dw, uint32_t dh, float *dst, uint32_t dstStride,
                                   size_t dstBatchStride, uint32_t sw, uint32_t sh, const float *src, uint32_t srcStride,
                                   size_t srcBatchStride)
{
    if (count == 0) return MVX_SUCCESS;
    if (dw != (sw + 1) / 2 || dh != (sh + 1) / 2 || !f32_plane_ok(dst, dw, dh, dstStride, dstBatchStride, count) ||
        !f32_plane_ok(src, sw, sh, srcStride, srcBatchStride, count) || count > 65535)
        return MVX_ERROR_INVALID_PARAMETERS;
    const F32Plane S{src, sw, sh, srcStride, srcBatchStride};
    hipLaunchKernelGGL(k_pyr_half_f32, f32_grid(dw, dh, count), dim3(WF_BX, WF_BY), 0, stream_of(stream), S, dst, dw, dh, dstStride,
                       dstBatchStride);
    MVX_LAUNCH_CHECK();
    return MVX_SUCCESS;
}

extern "C" int mvx_pyramid_half_f32_adjoint(mvx_stream_t stream, uint32_t count, uint32_t sw, uint32_t sh, float *gradSrc,
                                            uint32_t gsStride, size_t gsBatchStride, uint32_t dw, uint32_t dh, const float *gradDst,
                                            uint32_t gStride, size_t gBatchStride)
{
    if (count == 0) return MVX_SUCCESS;
    if (dw != (sw + 1) / 2 || dh != (sh + 1) / 2 || !f32_plane_ok(gradSrc, sw, sh, gsStride, gsBatchStride, count) ||
        !f32_plane_ok(gradDst, dw, dh, gStride, gBatchStride, count) || count > 65535)
        return MVX_ERROR_INVALID_PARAMETERS;
    const F32Plane G{gradDst, dw, dh, gStride, gBatchStride};
    hipLaunchKernelGGL(k_pyr_half_f32_adjoint, f32_grid(sw, sh, count), dim3(WF_BX, WF_BY), 0, stream_of(stream), G, gradSrc, sw, sh,
                       gsStride, gsBatchStride);
    MVX_LAUNCH_CHECK();
    return MVX_SUCCESS;
}
