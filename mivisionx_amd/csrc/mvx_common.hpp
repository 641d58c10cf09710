// mvx_common.hpp -- shared device helpers and launch plumbing for the gfx950 vision kernels.
//
// Image planes follow the reference's memory model (ago_util.cpp:187-200,1107):
//   * row stride is ALIGN16(width * bytes_per_pixel) and the padding bytes of every row are zero;
//   * neighbour reads are linear: byte index y*stride + x + dx, so the left neighbour of column 0
//     is the last byte of the previous row's stride, and indices outside [0, h*stride) read 0
//     (the reference's zero guard bytes).
// Kernels never write padding bytes, so a plane allocated zeroed keeps zero padding, with one
// exception copied from the reference: the warp and remap kernels store 4 pixels at a time and so
// write pixels w..ceil4(w)-1 of every row, as the reference's SSE loops do.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "../../include/mvx_kernels.h"

#define MVX_DEV __device__ __forceinline__

#define MVX_LAUNCH_CHECK()                                                     \
    do {                                                                       \
        hipError_t _e = hipGetLastError();                                     \
        if (_e != hipSuccess) return MVX_ERROR_LAUNCH;                         \
    } while (0)

namespace mvx {

// Device global memory reached through a generic pointer (one read out of a descriptor table rather
// than a kernel argument).  Accesses through the address_space(1) pointer issue as global_*
// instructions; generic ones issue as flat_*, which also count against the LDS wait counter, so an
// LDS read placed after a flat store waits until the store has reached memory.
template <class T> using gptr_t = __attribute__((address_space(1))) T *;
template <class T> MVX_DEV gptr_t<T> gmem(T *p) { return (gptr_t<T>)p; }

// Raw buffer access to one plane: a scalar resource holding the plane's base, the lane's column byte
// offset in a VGPR and the wave-uniform row offset in an SGPR, so a row's load or store needs no
// per-lane 64-bit address arithmetic.  Word 3 is the gfx9 raw-buffer format.  The record limit is the
// largest one, 2^32 - 1 bytes: planes of at most 32767 rows of at most 3 * 32767 + 15 bytes keep
// every row + column offset below it, so no access is ever clipped.
MVX_DEV __amdgpu_buffer_rsrc_t plane_rsrc(const void *base)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, (int)0xffffffffu, 0x00020000);
}

MVX_DEV uint32_t ld_u8(const uint8_t *p, long n, long i) { return (i < 0 || i >= n) ? 0u : (uint32_t)p[i]; }

// 16 bytes at a 16-byte aligned address inside the allocation
MVX_DEV uint4 ld16(const uint8_t *p) { return *reinterpret_cast<const uint4 *>(p); }

MVX_DEV uint32_t byte_of(const uint4 &v, int i)
{
    uint32_t w = (i < 4) ? v.x : (i < 8) ? v.y : (i < 12) ? v.z : v.w;
    return (w >> (8 * (i & 3))) & 0xffu;
}

MVX_DEV void unpack16(const uint4 &v, uint32_t (&o)[16])
{
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int i = 0; i < 16; i++) o[i] = (w[i >> 2] >> (8 * (i & 3))) & 0xffu;
}

MVX_DEV uint4 pack16(const uint32_t (&o)[16])
{
    uint32_t w[4];
#pragma unroll
    for (int k = 0; k < 4; k++)
        w[k] = (o[4 * k] & 0xff) | ((o[4 * k + 1] & 0xff) << 8) | ((o[4 * k + 2] & 0xff) << 16) | ((o[4 * k + 3] & 0xff) << 24);
    return make_uint4(w[0], w[1], w[2], w[3]);
}

// Store `n` (1..16) leading bytes of a 16-byte chunk; full chunks use one 128-bit store.
MVX_DEV void st16_partial(uint8_t *p, const uint32_t (&o)[16], int n)
{
    if (n >= 16) {
        *reinterpret_cast<uint4 *>(p) = pack16(o);
    } else {
        for (int i = 0; i < n; i++) p[i] = (uint8_t)o[i];
    }
}

// a / b for a divisor b fixed per launch, given y = RN(1/b) from recip_rn(b): the product a*y, its
// exact remainder r = a - b*(a*y) by FMA, and one FMA correction (Markstein's correction step).
// Where it replaces an IEEE division the replacement is checked exhaustively over every float a
// for the divisor used (tests/hip/probe_arith.hip); r == 0 returns a*y itself, which keeps the
// sign of a zero quotient.
MVX_DEV float div_by_const_rn(float a, float b, float y)
{
    const float q = a * y;
    const float r = fmaf(-q, b, a);
    return r == 0.0f ? q : fmaf(r, y, q);
}

// RN(1/b) in single precision: the double quotient is within half a double ulp of 1/b, and 1/b
// (b a float) is never a single-precision rounding midpoint, so rounding it to float is exact
inline float recip_rn(float b) { return (float)(1.0 / (double)b); }

inline hipStream_t stream_of(mvx_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

inline bool aligned16(const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

}  // namespace mvx
