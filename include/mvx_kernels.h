/*
 * mvx_kernels.h -- the C ABI of the MI355X (gfx950) vision-node kernels.
 *
 * Each entry point replaces one HIP launcher of the reference's AGO engine
 * (amd_openvx/openvx/hipvx/hip_host_decls.h, called from the ago_kernel_cmd_hip_execute branch of
 * the matching agoKernel_* function in amd_openvx/openvx/ago/ago_kernel_api.cpp) and keeps its
 * calling convention: a stream first, then the FULL output width/height, then device base
 * pointers with row strides in bytes, then scalars.  What changes is what the call computes:
 * every entry point here reproduces the reference CPU function (ago_haf_cpu*.cpp) named beside
 * it bit for bit -- row ranges, the linear-address border reads and all rounding -- which the
 * reference's own HIP kernels do not.
 *
 * Conventions
 *   - mvx_stream_t is a hipStream_t (0 = the default stream).  Calls are asynchronous on that
 *     stream: they enqueue work and return; nothing synchronises.
 *   - Buffers are device memory owned by the caller; no entry point allocates or frees, unless
 *     its comment says it uses a caller-provided workspace.
 *   - Image planes use the reference layout: stride = ALIGN16(width * bytes_per_pixel) (or any
 *     larger multiple of 16), base pointers 16-byte aligned, padding bytes zero.  Kernels never
 *     write padding bytes.
 *   - Return value: 0 on success, or a negative vx_status value (VX_ERROR_INVALID_PARAMETERS for
 *     bad dimensions/strides/alignment, VX_FAILURE when the launch itself fails).
 */
#ifndef MVX_KERNELS_H
#define MVX_KERNELS_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void *mvx_stream_t;

#define MVX_SUCCESS                   0
#define MVX_FAILURE                 (-1)   /* VX_FAILURE */
#define MVX_ERROR_NOT_IMPLEMENTED   (-2)   /* VX_ERROR_NOT_IMPLEMENTED */
#define MVX_ERROR_NO_MEMORY         (-8)   /* VX_ERROR_NO_MEMORY */
#define MVX_ERROR_INVALID_PARAMETERS (-10) /* VX_ERROR_INVALID_PARAMETERS */
#define MVX_ERROR_LAUNCH            MVX_FAILURE

/* Library version and build target ("gfx950"). */
const char *mvx_version(void);

/* ------------------------------------------------------------------ filters (3x3, U8) ----- */
/* Rows 1..h-2 are written; rows 0 and h-1 are left untouched, as the reference wrapper calls the
 * CPU function on height-2 rows (ago_kernel_api.cpp:15281-15282). */

/* replaces HipExec_Gaussian_U8_U8_3x3 (hipvx/hip_host_decls.h:649); semantics of
 * HafCpu_Gaussian_U8_U8_3x3 (ago_haf_cpu_filter.cpp:1518-1864) */
int mvx_Gaussian_U8_U8_3x3(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight,
                           uint8_t *pDst, uint32_t dstStride, const uint8_t *pSrc, uint32_t srcStride);

/* replaces HipExec_Box_U8_U8_3x3 (hip_host_decls.h:633); HafCpu_Box_U8_U8_3x3 (ago_haf_cpu_filter.cpp:34) */
int mvx_Box_U8_U8_3x3(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight,
                      uint8_t *pDst, uint32_t dstStride, const uint8_t *pSrc, uint32_t srcStride);

/* replaces HipExec_Median_U8_U8_3x3 (hip_host_decls.h:645); HafCpu_Median_U8_U8_3x3 (ago_haf_cpu_filter.cpp:3842) */
int mvx_Median_U8_U8_3x3(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight,
                         uint8_t *pDst, uint32_t dstStride, const uint8_t *pSrc, uint32_t srcStride);

/* replaces HipExec_Sobel_S16S16_U8_3x3_GXY (hip_host_decls.h:663), and with one of the outputs
 * NULL HipExec_Sobel_S16_U8_3x3_GX/_GY (:668/:672); HafCpu_Sobel_* (ago_haf_cpu_filter.cpp:1872,
 * :2047, :2237) */
int mvx_Sobel_S16S16_U8_3x3_GXY(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight,
                                int16_t *pDstGx, uint32_t dstGxStride, int16_t *pDstGy, uint32_t dstGyStride,
                                const uint8_t *pSrc, uint32_t srcStride);

/* ------------------------------------------------------------------ arithmetic ------------- */
/* replaces HipExec_Magnitude_S16_S16S16 (hip_host_decls.h:219); HafCpu_Magnitude_S16_S16S16
 * (ago_haf_cpu_arithmetic.cpp:4989-5097), all rows */
int mvx_Magnitude_S16_S16S16(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight,
                             int16_t *pDst, uint32_t dstStride, const int16_t *pGx, uint32_t gxStride,
                             const int16_t *pGy, uint32_t gyStride);

/* replaces HipExec_Phase_U8_S16S16 (hip_host_decls.h:224); HafCpu_Phase_U8_S16S16
 * (ago_haf_cpu_arithmetic.cpp:9592-9631), all rows */
int mvx_Phase_U8_S16S16(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight,
                        uint8_t *pDst, uint32_t dstStride, const int16_t *pGx, uint32_t gxStride,
                        const int16_t *pGy, uint32_t gyStride);

#ifdef __cplusplus
}
#endif
#endif
