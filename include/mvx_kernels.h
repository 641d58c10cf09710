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
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void *mvx_stream_t;

#define MVX_SUCCESS                   0
#define MVX_FAILURE                 (-1)   /* VX_FAILURE */
#define MVX_ERROR_NOT_IMPLEMENTED   (-2)   /* VX_ERROR_NOT_IMPLEMENTED */
#define MVX_ERROR_NOT_SUPPORTED     (-3)   /* VX_ERROR_NOT_SUPPORTED */
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

/* Merged Sobel -> Magnitude / Phase: the kernels the reference's graph optimizer substitutes for
 * Sobel_S16S16_U8_3x3_GXY + Magnitude_S16_S16S16 [+ Phase_U8_S16S16] when gx and gy are virtual and no
 * other node reads them (merge rules ago_drama_remove.cpp:155-182).  CPU-only in the reference
 * (ago_kernel_api.cpp:15958-16110 have no hip_execute branch), so there is no HipExec_* launcher to
 * replace; the argument order follows the HafCpu_* functions (output planes first).  Rows 1..h-2 are
 * written.  Magnitude: columns [0, w & ~15) round half to even, the last w mod 16 columns truncate. */
/* HafCpu_SobelMagnitudePhase_S16U8_U8_3x3 (ago_haf_cpu_filter.cpp:3972); pPhase may be NULL */
int mvx_SobelMagnitudePhase_S16U8_U8_3x3(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight,
                                         int16_t *pMag, uint32_t magStride, uint8_t *pPhase, uint32_t phaseStride,
                                         const uint8_t *pSrc, uint32_t srcStride);
/* HafCpu_SobelMagnitude_S16_U8_3x3 (ago_haf_cpu_filter.cpp:2468) */
int mvx_SobelMagnitude_S16_U8_3x3(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight,
                                  int16_t *pMag, uint32_t magStride, const uint8_t *pSrc, uint32_t srcStride);
/* HafCpu_SobelPhase_U8_U8_3x3 (ago_haf_cpu_filter.cpp:3946): the reference's Gx/Gy scratch rows are
 * ALIGN16(w) bytes apart for w int16 values, so pixel (x, y) takes the gradient of (x - ALIGN16(w)/2,
 * y + 1) for x >= ALIGN16(w)/2 on every row but the last; reproduced without a scratch buffer. */
int mvx_SobelPhase_U8_U8_3x3(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight,
                             uint8_t *pPhase, uint32_t phaseStride, const uint8_t *pSrc, uint32_t srcStride);

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

/* ------------------------------------------------------------------ colour / copy / scale -- */
/* ColorConvert RGB -> luma.  The reference's graph keeps only this kernel for RGB -> IYUV ->
 * ChannelExtract(Y) (ago_drama_divide.cpp:244-256 and the unused-output removal); it has no HIP
 * launcher there (CPU HafCpu_ColorConvert_Y_RGB, ago_haf_cpu_color_convert.cpp:2913-2990). */
int mvx_ColorConvert_Y_RGB(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight,
                           uint8_t *pDstY, uint32_t dstStride, const uint8_t *pSrcRGB, uint32_t srcStride);

/* replaces HipExec_ChannelCopy_U8_U8 (hip_host_decls.h:382); HafCpu_ChannelCopy_U8_U8
 * (ago_haf_cpu_ch_extract_combine.cpp:290); level 0 of a Gaussian pyramid */
int mvx_ChannelCopy_U8_U8(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight,
                          uint8_t *pDst, uint32_t dstStride, const uint8_t *pSrc, uint32_t srcStride);

/* replaces HipExec_ChannelExtract_U8_U16_Pos0 / _Pos1 (hip_host_decls.h:413, :417); HafCpu_ChannelExtract_U8_U16_Pos0 /
 * _Pos1 (ago_haf_cpu_ch_extract_combine.cpp:663-760): byte `pos` (0 or 1) of each 2-byte pixel. */
int mvx_ChannelExtract_U8_U16(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight, uint8_t *pDst, uint32_t dstStride,
                              const uint8_t *pSrc, uint32_t srcStride, uint32_t pos);
/* replaces HipExec_ChannelExtract_U8_U24_Pos0 / _Pos1 / _Pos2 (hip_host_decls.h:421, :425, :429);
 * HafCpu_ChannelExtract_U8_U24_Pos0..2 (ago_haf_cpu_ch_extract_combine.cpp:759 on): byte `pos` (0..2) of each RGB pixel. */
int mvx_ChannelExtract_U8_U24(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight, uint8_t *pDst, uint32_t dstStride,
                              const uint8_t *pSrc, uint32_t srcStride, uint32_t pos);

/* replace HipExec_ColorConvert_RGB_IYUV / _RGB_NV12 (hip_host_decls.h:523, :529); HafCpu_ColorConvert_RGB_IYUV /
 * _RGB_NV12 (ago_haf_cpu_color_convert.cpp:472-768): BT.709 with one chroma sample per 2x2 block, truncated
 * and clamped per channel.  Width and height must be even (-15 otherwise); pDst is packed RGB. */
int mvx_ColorConvert_RGB_IYUV(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight, uint8_t *pDst, uint32_t dstStride,
                              const uint8_t *pY, uint32_t yStride, const uint8_t *pU, uint32_t uStride, const uint8_t *pV,
                              uint32_t vStride);
int mvx_ColorConvert_RGB_NV12(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight, uint8_t *pDst, uint32_t dstStride,
                              const uint8_t *pY, uint32_t yStride, const uint8_t *pUV, uint32_t uvStride);

/* replaces HafCpu_ColorConvert_IYUV_RGB (ago_haf_cpu_color_convert.cpp:2492-2705; AGO kernel
 * COLOR_CONVERT_IYUV_RGB, ago_kernel_api.cpp:13447) and HipExec_ColorConvert_IYUV_RGB (hip_host_decls.h:565):
 * all three planes of IYUV from packed RGB, BT.709, chroma averaged over each 2x2 block with the reference's
 * rounding.  Width and height must be even (-15 otherwise); U and V are (w/2) x (h/2). */
int mvx_ColorConvert_IYUV_RGB(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight, uint8_t *pDstY, uint32_t yStride,
                              uint8_t *pDstU, uint32_t uStride, uint8_t *pDstV, uint32_t vStride, const uint8_t *pSrc,
                              uint32_t srcStride);

/* The rest of the ColorConvert family (ago_haf_cpu_color_convert.cpp), each replacing the HipExec_ launcher
 * of the same name (hip_host_decls.h:495-629) with the results of the HafCpu_ function of the same name.
 * Widths (and, for 4:2:0 data, heights) must be even; planes as in the conventions above; 4:2:0 chroma
 * planes are (w/2) x (h/2) samples (NV12 / NV21: interleaved pairs, 2 bytes per sample).
 *   YUV -> RGB / RGBX, BT.709 (:280-1732; the 4:2:2 scalar tails, columns w & ~7 .. w-1, group G's terms
 *   differently): */
int mvx_ColorConvert_RGB_UYVY(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight, uint8_t *pDst, uint32_t dstStride,
                              const uint8_t *pSrc, uint32_t srcStride);
int mvx_ColorConvert_RGB_YUYV(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight, uint8_t *pDst, uint32_t dstStride,
                              const uint8_t *pSrc, uint32_t srcStride);
int mvx_ColorConvert_RGBX_UYVY(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight, uint8_t *pDst, uint32_t dstStride,
                               const uint8_t *pSrc, uint32_t srcStride);
int mvx_ColorConvert_RGBX_YUYV(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight, uint8_t *pDst, uint32_t dstStride,
                               const uint8_t *pSrc, uint32_t srcStride);
int mvx_ColorConvert_RGB_NV21(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight, uint8_t *pDst, uint32_t dstStride,
                              const uint8_t *pSrcLuma, uint32_t lumaStride, const uint8_t *pSrcChroma, uint32_t chromaStride);
int mvx_ColorConvert_RGBX_NV12(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight, uint8_t *pDst, uint32_t dstStride,
                               const uint8_t *pSrcLuma, uint32_t lumaStride, const uint8_t *pSrcChroma, uint32_t chromaStride);
int mvx_ColorConvert_RGBX_NV21(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight, uint8_t *pDst, uint32_t dstStride,
                               const uint8_t *pSrcLuma, uint32_t lumaStride, const uint8_t *pSrcChroma, uint32_t chromaStride);
int mvx_ColorConvert_RGBX_IYUV(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight, uint8_t *pDst, uint32_t dstStride,
                               const uint8_t *pSrcY, uint32_t yStride, const uint8_t *pSrcU, uint32_t uStride, const uint8_t *pSrcV,
                               uint32_t vStride);
/*   RGB / RGBX -> NV12, IYUV, YUV4 (the merged kernels the reference's node-merge rules select,
 *   ago_drama_remove.cpp:38-100: NV12_RGB :2707, IYUV_RGBX :3868, NV12_RGBX :4078, YUV4_RGB :4282,
 *   YUV4_RGBX :3740): */
int mvx_ColorConvert_NV12_RGB(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight, uint8_t *pDstLuma, uint32_t lumaStride,
                              uint8_t *pDstChroma, uint32_t chromaStride, const uint8_t *pSrc, uint32_t srcStride);
int mvx_ColorConvert_NV12_RGBX(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight, uint8_t *pDstLuma, uint32_t lumaStride,
                               uint8_t *pDstChroma, uint32_t chromaStride, const uint8_t *pSrc, uint32_t srcStride);
int mvx_ColorConvert_IYUV_RGBX(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight, uint8_t *pDstY, uint32_t yStride,
                               uint8_t *pDstU, uint32_t uStride, uint8_t *pDstV, uint32_t vStride, const uint8_t *pSrc,
                               uint32_t srcStride);
int mvx_ColorConvert_YUV4_RGB(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight, uint8_t *pDstY, uint32_t yStride,
                              uint8_t *pDstU, uint32_t uStride, uint8_t *pDstV, uint32_t vStride, const uint8_t *pSrc,
                              uint32_t srcStride);
int mvx_ColorConvert_YUV4_RGBX(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight, uint8_t *pDstY, uint32_t yStride,
                               uint8_t *pDstU, uint32_t uStride, uint8_t *pDstV, uint32_t vStride, const uint8_t *pSrc,
                               uint32_t srcStride);
/*   RGBX -> luma alone (ColorConvert_Y_RGBX :3367, what the reference keeps when only Y is read): */
int mvx_ColorConvert_Y_RGBX(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight, uint8_t *pDstY, uint32_t yStride,
                            const uint8_t *pSrc, uint32_t srcStride);
/*   RGB <-> RGBX (:2330, :2399): */
int mvx_ColorConvert_RGBX_RGB(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight, uint8_t *pDst, uint32_t dstStride,
                              const uint8_t *pSrc, uint32_t srcStride);
int mvx_ColorConvert_RGB_RGBX(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight, uint8_t *pDst, uint32_t dstStride,
                              const uint8_t *pSrc, uint32_t srcStride);
/*   YUV -> YUV byte moves: 4:2:2 -> 4:2:0 with the chroma rows' floor average (:55, :1733, :1958, :2144);
 *   UV12_IUV interleaves and IUV_UV12 de-interleaves (cw x ch = the chroma plane's samples; NV21 callers
 *   swap the U / V pointers, as the reference's divider does); UV_UV12 and ScaleUp2x2 replicate each
 *   chroma sample over a 2x2 block of full-size planes (dstWidth x dstHeight = the full size). */
int mvx_FormatConvert_IYUV_UYVY(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight, uint8_t *pDstY, uint32_t yStride,
                                uint8_t *pDstU, uint32_t uStride, uint8_t *pDstV, uint32_t vStride, const uint8_t *pSrc,
                                uint32_t srcStride);
int mvx_FormatConvert_IYUV_YUYV(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight, uint8_t *pDstY, uint32_t yStride,
                                uint8_t *pDstU, uint32_t uStride, uint8_t *pDstV, uint32_t vStride, const uint8_t *pSrc,
                                uint32_t srcStride);
int mvx_FormatConvert_NV12_UYVY(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight, uint8_t *pDstLuma, uint32_t lumaStride,
                                uint8_t *pDstChroma, uint32_t chromaStride, const uint8_t *pSrc, uint32_t srcStride);
int mvx_FormatConvert_NV12_YUYV(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight, uint8_t *pDstLuma, uint32_t lumaStride,
                                uint8_t *pDstChroma, uint32_t chromaStride, const uint8_t *pSrc, uint32_t srcStride);
int mvx_FormatConvert_IUV_UV12(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight, uint8_t *pDstU, uint32_t uStride,
                               uint8_t *pDstV, uint32_t vStride, const uint8_t *pSrcChroma, uint32_t chromaStride);
int mvx_FormatConvert_UV12_IUV(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight, uint8_t *pDstChroma, uint32_t chromaStride,
                               const uint8_t *pSrcU, uint32_t uStride, const uint8_t *pSrcV, uint32_t vStride);
int mvx_FormatConvert_UV_UV12(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight, uint8_t *pDstU, uint32_t uStride,
                              uint8_t *pDstV, uint32_t vStride, const uint8_t *pSrcChroma, uint32_t chromaStride);
int mvx_ScaleUp2x2_U8_U8(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight, uint8_t *pDst, uint32_t dstStride,
                         const uint8_t *pSrc, uint32_t srcStride);

/* replaces HipExec_ScaleImage_U8_U8_Bilinear (hip_host_decls.h:728); HafCpu_ScaleImage_U8_U8_Bilinear
 * (ago_haf_cpu_geometric.cpp:1977-2105) with the scale matrix of ago_kernel_api.cpp:19897-19916.
 * dstStride must be a multiple of 16 (8-byte stores). */
int mvx_ScaleImage_U8_U8_Bilinear(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight,
                                  uint8_t *pDst, uint32_t dstStride, uint32_t srcWidth, uint32_t srcHeight,
                                  const uint8_t *pSrc, uint32_t srcStride);

/* replaces HipExec_ScaleImage_U8_U8_Nearest (hip_host_decls.h:723); HafCpu_ScaleImage_U8_U8_Nearest
 * (ago_haf_cpu_geometric.cpp:1874-1976) with offsets ratio/2 (ago_kernel_api.cpp:19790-19793).
 * Sizes up to 65535 (the reference's 16-bit maps); dstStride a multiple of 16. */
int mvx_ScaleImage_U8_U8_Nearest(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight,
                                 uint8_t *pDst, uint32_t dstStride, uint32_t srcWidth, uint32_t srcHeight,
                                 const uint8_t *pSrc, uint32_t srcStride);

/* replace HipExec_ScaleImage_U8_U8_Bilinear_Replicate / _Constant (hip_host_decls.h:733, :738);
 * HafCpu_ScaleImage_U8_U8_Bilinear_Replicate (ago_haf_cpu_geometric.cpp:2108-2258) and _Constant
 * (:2260-2391), scale matrix of ago_kernel_api.cpp:20017-20020 / :20147-20150, including the
 * reference's split between its two-stage-rounded SIMD columns and single-rounded tail columns.
 * srcWidth >= 2; dstStride a multiple of 16. */
int mvx_ScaleImage_U8_U8_Bilinear_Replicate(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight,
                                            uint8_t *pDst, uint32_t dstStride, uint32_t srcWidth, uint32_t srcHeight,
                                            const uint8_t *pSrc, uint32_t srcStride);
int mvx_ScaleImage_U8_U8_Bilinear_Constant(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight,
                                           uint8_t *pDst, uint32_t dstStride, uint32_t srcWidth, uint32_t srcHeight,
                                           const uint8_t *pSrc, uint32_t srcStride, uint8_t borderValue);

/* replaces HipExec_ScaleImage_U8_U8_Area (hip_host_decls.h:744); HafCpu_ScaleImage_U8_U8_Area
 * (ago_haf_cpu_geometric.cpp:2491-2617), offsets -1/2 (ago_kernel_api.cpp:20276-20277), as a
 * reference graph runs it after the optimizer's rewrite (ago_drama_remove.cpp:1270-1311): a plane
 * copy at the same size, nearest neighbour unless both dimensions shrink, and the 2x2 average
 * ((a+b+c+d+2)>>2, columns below dstWidth & ~7 only; dstWidth >= 8) at an exact 2x.  Other
 * shrinking ratios return MVX_ERROR_NOT_SUPPORTED: the reference's generic path stores its row sums
 * past the end of its node buffer, so its result is not defined. */
int mvx_ScaleImage_U8_U8_Area(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight,
                              uint8_t *pDst, uint32_t dstStride, uint32_t srcWidth, uint32_t srcHeight,
                              const uint8_t *pSrc, uint32_t srcStride);

/* replaces HipExec_ScaleGaussianHalf_U8_U8_3x3 (hip_host_decls.h:676); HafCpu_ScaleGaussianHalf_U8_U8_3x3
 * (ago_haf_cpu_geometric.cpp:2772-2829) as its wrapper calls it (ago_kernel_api.cpp:15341-15352): output
 * rows 1..dstHeight-2, row y centred on source row 2y, [1 2 1]^2 >> 4 with a zero left neighbour at
 * column 0.  Returns -15 (VX_ERROR_INVALID_DIMENSION) for dstHeight == 1 or an output larger than half
 * the source rounded up; dstStride a multiple of 8. */
int mvx_ScaleGaussianHalf_U8_U8_3x3(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight,
                                    uint8_t *pDst, uint32_t dstStride, uint32_t srcWidth, uint32_t srcHeight,
                                    const uint8_t *pSrc, uint32_t srcStride);

/* replaces HipExec_ScaleGaussianHalf_U8_U8_5x5 (hip_host_decls.h:681); HafCpu_ScaleGaussianHalf_U8_U8_5x5
 * (ago_haf_cpu_pyramid.cpp:94-305) as called by ago_kernel_api.cpp:15420-15436 (rows 1..h-2, parity
 * of the source size picks the sampled row/column).  Returns -15 (VX_ERROR_INVALID_DIMENSION) when the
 * source is below 5 or the destination below 3 in either dimension, as the reference wrapper does. */
int mvx_ScaleGaussianHalf_U8_U8_5x5(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight,
                                    uint8_t *pDst, uint32_t dstStride, uint32_t srcWidth, uint32_t srcHeight,
                                    const uint8_t *pSrc, uint32_t srcStride);

/* ------------------------------------------------------------------ Canny (3x3, L1) -------- */
/* replaces HipExec_CannySobel_U16_U8_3x3_L1NORM (hip_host_decls.h:828); HafCpu_CannySobel_U16_U8_3x3_L1NORM
 * (ago_haf_cpu_canny.cpp:44-136), including its column-0 values and, below 16 pixels of width, its
 * write of the U16 slot before each row. */
int mvx_CannySobel_U16_U8_3x3_L1NORM(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight,
                                     uint16_t *pDst, uint32_t dstStride, const uint8_t *pSrc, uint32_t srcStride);

/* replace HipExec_CannySobel_U16_U8_5x5_L1NORM, _7x7_L1NORM, _3x3_L2NORM, _5x5_L2NORM, _7x7_L2NORM
 * (hip_host_decls.h:832, 836, 840, 844, 848); HafCpu_CannySobel_U16_U8_* (ago_haf_cpu_canny.cpp:145-427,
 * 646-1056): rows R..h-R-1 (R = gradient/2), ((magnitude << 2) | sector) in 16 bits, the reference's
 * 16-bit wrapping/saturating lane arithmetic, its scratch-row reads at the ends of the SSE columns and
 * its scalar loop from column (w & ~15) - 1 (from -1 below 16 pixels of width, writing the slot before
 * each row).  The suppression stage is mvx_CannySuppThreshold_U8XY_U16_3x3 for every size (for
 * gradient 7 with the thresholds divided by 4, as the reference's 7x7 stage does). */
int mvx_CannySobel_U16_U8_5x5_L1NORM(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight, uint16_t *pDst,
                                     uint32_t dstStride, const uint8_t *pSrc, uint32_t srcStride);
int mvx_CannySobel_U16_U8_7x7_L1NORM(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight, uint16_t *pDst,
                                     uint32_t dstStride, const uint8_t *pSrc, uint32_t srcStride);
int mvx_CannySobel_U16_U8_3x3_L2NORM(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight, uint16_t *pDst,
                                     uint32_t dstStride, const uint8_t *pSrc, uint32_t srcStride);
int mvx_CannySobel_U16_U8_5x5_L2NORM(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight, uint16_t *pDst,
                                     uint32_t dstStride, const uint8_t *pSrc, uint32_t srcStride);
int mvx_CannySobel_U16_U8_7x7_L2NORM(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight, uint16_t *pDst,
                                     uint32_t dstStride, const uint8_t *pSrc, uint32_t srcStride);

/* replaces HipExec_CannySuppThreshold_U8XY_U16_3x3 (hip_host_decls.h:852); HafCpu_CannySuppThreshold_U8XY_U16_3x3
 * (ago_haf_cpu_canny.cpp:549-596).  Writes 0/127/255 over [1,w-2]x[1,h-2].  xyStack is the
 * ago-canny-stack buffer (capacity >= w*h items of 4 bytes, ago_drama_divide.cpp:1633); it is
 * accepted for interface parity and used as workspace by mvx_CannyEdgeTrace_U8_U8XY. */
int mvx_CannySuppThreshold_U8XY_U16_3x3(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight,
                                        uint8_t *pDst, uint32_t dstStride, const uint16_t *pMag, uint32_t magStride,
                                        void *xyStack, uint32_t capacityOfXY, uint32_t hystLower, uint32_t hystUpper);

/* Hysteresis of HafCpu_CannyEdgeTrace_U8_U8XY (ago_haf_cpu_canny.cpp:598-644; CPU-only in the
 * reference, ago_kernel_list.cpp:486): 127 pixels 8-connected through 127s to a 255 become 255, all
 * other 127s become 0, over [1,w-2]x[1,h-2].  In place on pImg.  xyStack: >= w*h * 4 bytes of device
 * workspace at a 4-byte aligned address (its contents are not read; otherwise -10); the stack of a
 * large enough image holds the tiled trace's tables, else the per-pixel union-find's labels. */
int mvx_CannyEdgeTrace_U8_U8XY(mvx_stream_t stream, uint32_t width, uint32_t height, uint8_t *pImg,
                               uint32_t stride, void *xyStack, uint32_t capacityOfXY);

/* ------------------------------------------------------------------ FAST corners ----------- */
/* replace HipExec_FastCorners_XY_U8_NoSupression / _Supression (hip_host_decls.h:858, :863), with the
 * results of HafCpu_FastCorners_XY_U8_NoSupression / _Supression (ago_haf_cpu_fast_corners.cpp:414-790):
 * FAST-9 on the radius-3 circle with t = (short)floorf(threshold) and the reference's early exit; without
 * suppression candidates at every x of rows 3..h-4 with strength = threshold, with suppression x in
 * [3, w-4] and the bisected strength, then the 3x3 test (>= upper and left, > right and lower).
 * Keypoints (vx_keypoint_t, 28 bytes) in raster order up to `capacity`; *dstCount (device) receives
 * min(count, capacity) and *numCorners (device, optional) the uncapped count.  Images narrower or
 * lower than 7 return -15.  The workspace holds mvx_FastCorners_workspace_size(w, h) bytes. */
size_t mvx_FastCorners_workspace_size(uint32_t width, uint32_t height);
int mvx_FastCorners_XY_U8_NoSupression(mvx_stream_t stream, uint32_t capacity, void *dstCorners, uint32_t *dstCount,
                                       uint64_t *numCorners, uint32_t width, uint32_t height, const uint8_t *pSrc,
                                       uint32_t srcStride, float threshold, void *workspace, size_t workspaceSize);
int mvx_FastCorners_XY_U8_Supression(mvx_stream_t stream, uint32_t capacity, void *dstCorners, uint32_t *dstCount,
                                     uint64_t *numCorners, uint32_t width, uint32_t height, const uint8_t *pSrc,
                                     uint32_t srcStride, float threshold, void *workspace, size_t workspaceSize);

/* ------------------------------------------------------------------ Harris (3x3, 3x3) ------ */
/* replaces HipExec_HarrisSobel_HG3_U8_3x3 (hip_host_decls.h:868); HafCpu_HarrisSobel_HG3_U8_3x3
 * (ago_haf_cpu_harris.cpp:85-169).  HG3 is an F32x3 plane (stride >= ALIGN16(12w)); rows are laid
 * out exactly as the reference writes them (row 1 at byte `stride`, then stride/12-pixel steps). */
int mvx_HarrisSobel_HG3_U8_3x3(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight,
                               float *pDstGxy, uint32_t dstStride, const uint8_t *pSrc, uint32_t srcStride);

/* replaces HipExec_HarrisScore_HVC_HG3_3x3 (hip_host_decls.h:880); HafCpu_HarrisScore_HVC_HG3_3x3
 * (ago_haf_cpu_harris.cpp:596-659); normFactor as computed by ago_kernel_api.cpp:16941-16942. */
int mvx_HarrisScore_HVC_HG3_3x3(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight,
                                float *pDstVc, uint32_t dstStride, const float *pSrcGxy, uint32_t srcStride,
                                float sensitivity, float strengthThreshold, float normFactor);

/* replaces HipExec_NonMaxSupp_XY_ANY_3x3 (hip_host_decls.h:898); HafCpu_NonMaxSupp_XY_ANY_3x3
 * (ago_haf_cpu_harris.cpp:882-921).  Items are {int16 x, int16 y, float s} (ago_keypoint_xys_t) in
 * raster order, at most `capacity`; *pDstCount (device) receives the item count.  The workspace
 * (device, mvx_NonMaxSupp_XY_ANY_3x3_workspace_size bytes) holds per-block counts and offsets. */
size_t mvx_NonMaxSupp_XY_ANY_3x3_workspace_size(uint32_t srcWidth, uint32_t srcHeight);
int mvx_NonMaxSupp_XY_ANY_3x3(mvx_stream_t stream, uint32_t capacity, void *pDstList, uint32_t *pDstCount,
                              uint32_t srcWidth, uint32_t srcHeight, const float *pSrc, uint32_t srcStride,
                              void *workspace, size_t workspaceSize);

/* No counterpart in the reference.  For callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc): mvx_NonMaxSupp_XY_ANY_3x3
 * without the workspace argument.  Its counts and offsets live in a library-owned device array of
 * MVX_NONMAXSUPP_NOALLOC_SCRATCH_BYTES per GPU (held until the process ends), enough for images
 * whose mvx_NonMaxSupp_XY_ANY_3x3_workspace_size is at most that (e.g. 8192 x 4096); larger images
 * return MVX_ERROR_NO_MEMORY and launch nothing.  Nothing is allocated or freed.  Every call on one
 * GPU shares that GPU's array, so calls that overlap on one GPU (other streams, other threads) race. */
#define MVX_NONMAXSUPP_NOALLOC_SCRATCH_BYTES (1u << 20)
int mvx_NonMaxSupp_XY_ANY_3x3_NoAlloc(mvx_stream_t stream, uint32_t capacity, void *pDstList, uint32_t *pDstCount,
                                      uint32_t srcWidth, uint32_t srcHeight, const float *pSrc, uint32_t srcStride);

/* replace HipExec_HarrisSobel_HG3_U8_5x5 / _7x7 (hip_host_decls.h:872, :876); HafCpu_HarrisSobel_HG3_U8_5x5
 * / _7x7 (ago_haf_cpu_harris.cpp:333-594): rows R..h-R-1 (R = 2, 3), written from byte R*stride as packed
 * rows of w pixels, as the reference writes them. */
int mvx_HarrisSobel_HG3_U8_5x5(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight,
                               float *pDstGxy, uint32_t dstStride, const uint8_t *pSrc, uint32_t srcStride);
int mvx_HarrisSobel_HG3_U8_7x7(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight,
                               float *pDstGxy, uint32_t dstStride, const uint8_t *pSrc, uint32_t srcStride);
/* replace HipExec_HarrisScore_HVC_HG3_5x5 / _7x7 (hip_host_decls.h:886, :892); HafCpu_HarrisScore_HVC_HG3_5x5
 * / _7x7 (ago_haf_cpu_harris.cpp:661-802), including the 7x7 form's one-column left shift of its
 * outputs; normFactor = (255 * 2^(g-1) * block)^4 (ago_kernel_api.cpp:17012-17013, :17083-17084).  Planes
 * narrower or lower than the block return -15 (the reference's border stores would overrun them). */
int mvx_HarrisScore_HVC_HG3_5x5(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight,
                                float *pDstVc, uint32_t dstStride, const float *pSrcGxy, uint32_t srcStride,
                                float sensitivity, float strengthThreshold, float normFactor);
int mvx_HarrisScore_HVC_HG3_7x7(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight,
                                float *pDstVc, uint32_t dstStride, const float *pSrcGxy, uint32_t srcStride,
                                float sensitivity, float strengthThreshold, float normFactor);

/* Sort and pick of HafCpu_HarrisMergeSortAndPick_XY_XYS (ago_haf_cpu_harris.cpp:923-1005; CPU-only in
 * the reference, ago_kernel_list.cpp:483), grid of ago_kernel_api.cpp:20808-20822.  Reads
 * *pSrcCount (device) items of pSrcList (at most maxList); writes vx_keypoint_t items (28 bytes) to
 * pDstCorners in descending key order up to `capacity`, *pDstNumItems = min(count, capacity) and
 * *pDstCount = count (vx_size, uncapped).  width/height are the image size (below 32768).
 * The workspace (mvx_HarrisMergeSortAndPick_XY_XYS_workspace_size(maxList) bytes) must be prepared
 * once by mvx_HarrisMergeSortAndPick_workspace_init; every call leaves it ready for the next. */
size_t mvx_HarrisMergeSortAndPick_XY_XYS_workspace_size(uint32_t maxList);
int mvx_HarrisMergeSortAndPick_workspace_init(mvx_stream_t stream, uint32_t maxList, void *workspace,
                                              size_t workspaceSize);
int mvx_HarrisMergeSortAndPick_XY_XYS(mvx_stream_t stream, uint32_t capacity, void *pDstCorners,
                                      uint32_t *pDstNumItems, uint64_t *pDstCount, const void *pSrcList,
                                      const uint32_t *pSrcCount, uint32_t maxList, float minDistance,
                                      uint32_t width, uint32_t height, void *workspace, size_t workspaceSize);

/* ------------------------------------------------------------------ warps and remap -------- */
/* Bilinear U8 warps and remap.  Rows are written up to ceil(w/4)*4 pixels (the reference's 4-pixel
 * stores; padding pixels hold what the reference leaves there: column 0's value for the warps, the
 * next map entries' for remap), so dstStride must be >= ceil(w/4)*4 and a multiple of 4; the source is read at linear
 * byte addresses with reads outside [0, srcHeight*srcStride) returning 0.  `matrix` is host memory
 * holding the VX matrix as stored: 6 floats (2 columns x 3 rows) for affine, 9 for perspective. */

/* replaces HipExec_WarpAffine_U8_U8_Bilinear (hip_host_decls.h:762); HafCpu_WarpAffine_U8_U8_Bilinear
 * (ago_haf_cpu_geometric.cpp:825-1099), border UNDEFINED */
int mvx_WarpAffine_U8_U8_Bilinear(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight, uint8_t *pDst,
                                  uint32_t dstStride, uint32_t srcWidth, uint32_t srcHeight, const uint8_t *pSrc,
                                  uint32_t srcStride, const float *matrix);
/* replaces HipExec_WarpAffine_U8_U8_Bilinear_Constant (hip_host_decls.h:768);
 * HafCpu_WarpAffine_U8_U8_Bilinear_Constant (ago_haf_cpu_geometric.cpp:1100-1370) */
int mvx_WarpAffine_U8_U8_Bilinear_Constant(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight, uint8_t *pDst,
                                           uint32_t dstStride, uint32_t srcWidth, uint32_t srcHeight, const uint8_t *pSrc,
                                           uint32_t srcStride, const float *matrix, uint8_t borderValue);
/* replaces HipExec_WarpPerspective_U8_U8_Bilinear (hip_host_decls.h:788); HafCpu_WarpPerspective_U8_U8_Bilinear
 * (ago_haf_cpu_geometric.cpp:1560-1577) */
int mvx_WarpPerspective_U8_U8_Bilinear(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight, uint8_t *pDst,
                                       uint32_t dstStride, uint32_t srcWidth, uint32_t srcHeight, const uint8_t *pSrc,
                                       uint32_t srcStride, const float *matrix);
/* replaces HipExec_WarpPerspective_U8_U8_Bilinear_Constant (hip_host_decls.h:794);
 * HafCpu_WarpPerspective_U8_U8_Bilinear_Constant (ago_haf_cpu_geometric.cpp:1579-1873) */
int mvx_WarpPerspective_U8_U8_Bilinear_Constant(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight, uint8_t *pDst,
                                                uint32_t dstStride, uint32_t srcWidth, uint32_t srcHeight,
                                                const uint8_t *pSrc, uint32_t srcStride, const float *matrix,
                                                uint8_t borderValue);
/* replaces HipExec_Remap_U8_U8_Bilinear (hip_host_decls.h:813); HafCpu_Remap_U8_U8_Bilinear
 * (ago_haf_cpu_geometric.cpp:249-370).  pMap: device table of {u16 x, u16 y} in 13.3 fixed point,
 * (0xffff, 0xffff) for points outside the source (mvx_Remap_FixedMap_F32 builds it), mapStride
 * bytes per row (the reference uses dstWidth*4).  srcImageBufferSize: bytes of the source buffer, as
 * the reference passes it; MVX_ERROR_INVALID_PARAMETERS when it does not cover
 * (srcHeight - 1) * srcStride + srcWidth bytes. */
int mvx_Remap_U8_U8_Bilinear(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight, uint8_t *pDst, uint32_t dstStride,
                             uint32_t srcWidth, uint32_t srcHeight, const uint8_t *pSrc, uint32_t srcStride,
                             uint32_t srcImageBufferSize, const void *pMap, uint32_t mapStride);
/* replaces HipExec_Remap_U8_U8_Bilinear_Constant (hip_host_decls.h:819); HafCpu_Remap_U8_U8_Bilinear_Constant
 * (ago_haf_cpu_geometric.cpp:372-504) */
int mvx_Remap_U8_U8_Bilinear_Constant(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight, uint8_t *pDst,
                                      uint32_t dstStride, uint32_t srcWidth, uint32_t srcHeight, const uint8_t *pSrc,
                                      uint32_t srcStride, const void *pMap, uint32_t mapStride, uint8_t borderValue);
/* Nearest-neighbour forms: the same conventions (the nearest remap writes no padding). */
/* replaces HipExec_WarpAffine_U8_U8_Nearest (hip_host_decls.h:749); HafCpu_WarpAffine_U8_U8_Nearest
 * (ago_haf_cpu_geometric.cpp:508-656): truncated coordinates, the corner test choosing the masked
 * loop (0 outside).  The reference's column tables fault on widths that are not a multiple of 4;
 * here every width is evaluated by the same formula. */
int mvx_WarpAffine_U8_U8_Nearest(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight, uint8_t *pDst,
                                 uint32_t dstStride, uint32_t srcWidth, uint32_t srcHeight, const uint8_t *pSrc,
                                 uint32_t srcStride, const float *matrix);
/* replaces HipExec_WarpAffine_U8_U8_Nearest_Constant (hip_host_decls.h:755);
 * HafCpu_WarpAffine_U8_U8_Nearest_Constant (ago_haf_cpu_geometric.cpp:658-823).  validRect: host
 * {start_x, start_y, end_x, end_y} of the source's valid region (the reference's rect_valid), NULL for
 * the whole image. */
int mvx_WarpAffine_U8_U8_Nearest_Constant(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight, uint8_t *pDst,
                                          uint32_t dstStride, uint32_t srcWidth, uint32_t srcHeight, const uint8_t *pSrc,
                                          uint32_t srcStride, const float *matrix, uint8_t borderValue,
                                          const uint32_t *validRect);
/* replaces HipExec_WarpPerspective_U8_U8_Nearest (hip_host_decls.h:775); HafCpu_WarpPerspective_U8_U8_Nearest
 * (ago_haf_cpu_geometric.cpp:1372-1385) */
int mvx_WarpPerspective_U8_U8_Nearest(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight, uint8_t *pDst,
                                      uint32_t dstStride, uint32_t srcWidth, uint32_t srcHeight, const uint8_t *pSrc,
                                      uint32_t srcStride, const float *matrix);
/* replaces HipExec_WarpPerspective_U8_U8_Nearest_Constant (hip_host_decls.h:781);
 * HafCpu_WarpPerspective_U8_U8_Nearest_Constant (ago_haf_cpu_geometric.cpp:1406-1558) */
int mvx_WarpPerspective_U8_U8_Nearest_Constant(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight, uint8_t *pDst,
                                               uint32_t dstStride, uint32_t srcWidth, uint32_t srcHeight,
                                               const uint8_t *pSrc, uint32_t srcStride, const float *matrix,
                                               uint8_t borderValue);
/* replaces HipExec_Remap_U8_U8_Nearest (hip_host_decls.h:801); HafCpu_Remap_U8_U8_Nearest
 * (ago_haf_cpu_geometric.cpp:72-154), including its different reading of (0xffff) lanes in the last
 * dstWidth % 4 columns */
int mvx_Remap_U8_U8_Nearest(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight, uint8_t *pDst, uint32_t dstStride,
                            uint32_t srcWidth, uint32_t srcHeight, const uint8_t *pSrc, uint32_t srcStride,
                            uint32_t srcImageBufferSize, const void *pMap, uint32_t mapStride);
/* replaces HipExec_Remap_U8_U8_Nearest_Constant (hip_host_decls.h:807); HafCpu_Remap_U8_U8_Nearest_Constant
 * (ago_haf_cpu_geometric.cpp:156-247) */
int mvx_Remap_U8_U8_Nearest_Constant(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight, uint8_t *pDst,
                                     uint32_t dstStride, uint32_t srcWidth, uint32_t srcHeight, const uint8_t *pSrc,
                                     uint32_t srcStride, uint32_t srcImageBufferSize, const void *pMap, uint32_t mapStride,
                                     uint8_t borderValue);
/* RGB (packed U24) bilinear remap for the stitch path.  The reference registers REMAP_U24_U24_BILINEAR
 * with an empty execute (ago_kernel_api.cpp:18535-18539); here each channel equals
 * mvx_Remap_U8_U8_Bilinear[_Constant] of that channel over the same 13.3 table (valid entries only
 * read inside the image, so the layout does not matter).  Exactly dstWidth pixels per row. */
int mvx_Remap_U24_U24_Bilinear(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight, uint8_t *pDst, uint32_t dstStride,
                               uint32_t srcWidth, uint32_t srcHeight, const uint8_t *pSrc, uint32_t srcStride,
                               const void *pMap, uint32_t mapStride);
int mvx_Remap_U24_U24_Bilinear_Constant(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight, uint8_t *pDst,
                                        uint32_t dstStride, uint32_t srcWidth, uint32_t srcHeight, const uint8_t *pSrc,
                                        uint32_t srcStride, const void *pMap, uint32_t mapStride, uint8_t borderValue);
/* Stitch merge of up to MVX_STITCH_MAX_CAMERAS remapped RGB camera images with per-camera U8 weight
 * planes (the role of Loom's merge, amd_loomsl/kernels/merge.cpp:223-290, OpenCL-only in the
 * reference): per channel the in-order float sum of (w_k / 255) * p_k, rounded to nearest and
 * saturated.  pImages / pWeights / strides are host arrays of device pointers and pitches. */
#define MVX_STITCH_MAX_CAMERAS 16
int mvx_StitchMerge_RGB(mvx_stream_t stream, uint32_t width, uint32_t height, uint8_t *pDst, uint32_t dstStride, uint32_t count,
                        const uint8_t *const *pImages, const uint32_t *imageStrides, const uint8_t *const *pWeights,
                        const uint32_t *weightStrides);
/* Loom's merge with its camera-selection maps (amd_loomsl/kernels/merge.cpp:176-385; OpenCL code
 * generation only in the reference, so there is no CPU or HIP form to replace).  Per group of 8 output
 * pixels camId0 (U8, one per group) holds a camera id (< 31: copy that camera), 31 (no camera: the
 * pixels are not written) or a blend code (> 31: the cameras listed as 5-bit ids in camId1 / camId2, U16
 * per group; 2 cameras up to 128, then one more for each step above 128), weighted by
 * weight * 0.003922f (the factor Loom's generated source prints with "%f"), summed in order and packed to
 * U8 with round-to-nearest-even and saturation.  Camera c's remapped RGB image and U8 weights are rows
 * c*height .. c*height+height-1 of the stacked planes (Loom stacks RGBX images the same way).  Ids at or
 * above `count` are treated as 31. */
int mvx_StitchMerge_Loom_RGB(mvx_stream_t stream, uint32_t width, uint32_t height, uint8_t *pDst, uint32_t dstStride,
                             uint32_t count, const uint8_t *pStacked, uint32_t stackedStride, const uint8_t *pWeights,
                             uint32_t weightStride, const uint8_t *pCamId0, uint32_t camId0Stride, const uint16_t *pCamId1,
                             uint32_t camId1Stride, const uint16_t *pCamId2, uint32_t camId2Stride);
/* The same stitch in one pass from the camera images themselves: camera c is sampled through its 13.3
 * remap table (rows c*height .. of pMaps, mvx_Remap_FixedMap_F32 of its F32 map) exactly as
 * mvx_Remap_U24_U24_Bilinear_Constant (border 0) would remap it, only where the selection needs it, so
 * the remapped frames are never written.  pCameras / cameraStrides: host arrays of `count` device
 * pointers and pitches of srcWidth x srcHeight RGB images.  Same bytes as the remaps followed by
 * mvx_StitchMerge_Loom_RGB. */
int mvx_Stitch_Loom_RGB(mvx_stream_t stream, uint32_t width, uint32_t height, uint8_t *pDst, uint32_t dstStride, uint32_t count,
                        const uint8_t *const *pCameras, const uint32_t *cameraStrides, uint32_t srcWidth, uint32_t srcHeight,
                        const void *pMaps, uint32_t mapStride, const uint8_t *pWeights, uint32_t weightStride,
                        const uint8_t *pCamId0, uint32_t camId0Stride, const uint16_t *pCamId1, uint32_t camId1Stride,
                        const uint16_t *pCamId2, uint32_t camId2Stride);
/* The fixed-point remap table of vxSetRemapPoint (vx_api.cpp:7865-7894, 3 fractional bits) from a
 * device float table of {x, y} pairs, on the GPU.  Sources must be below 8192 pixels in each
 * dimension (the reference uses fewer fractional bits beyond, ago_util.cpp:2450-2459). */
int mvx_Remap_FixedMap_F32(mvx_stream_t stream, uint32_t dstWidth, uint32_t dstHeight, void *pFixedMap, uint32_t fixedStride,
                           const float *pFloatMap, uint32_t floatStride, uint32_t srcWidth, uint32_t srcHeight);

/* ------------------------------------------------------------------ optical flow ----------- */
/* One level of an image pyramid in device memory. */
typedef struct mvx_image_level {
    const uint8_t *ptr;
    uint32_t width, height, stride;
} mvx_image_level_t;

/* Pyramidal Lucas-Kanade of HafCpu_OpticalFlowPyrLK_XY_XY_Generic (ago_haf_cpu_opticalflow.cpp:153-546),
 * as called by agoKernel_OpticalFlowPyrLK_XY_XY (ago_kernel_api.cpp:20587-20650); CPU-only in the
 * reference (ago_kernel_api.cpp:20642-20646).  oldPyramid / newPyramid: host arrays of `levels`
 * U8 levels (level 0 first; equal sizes; widths >= 8).  Keypoint arrays are device vx_keypoint_t
 * (28 bytes); *pOldCount (device) items of pOldKeypoints (and pEstimates when useInitialEstimate)
 * are tracked, at most `capacity`; *pNewCount (device, may be NULL) receives the same count.
 * termination: VX_TERM_CRITERIA_ITERATIONS / EPSILON / BOTH (EPSILON alone is capped at 10000
 * iterations; the reference loops without bound).  windowDimension 3..15.
 * The workspace (mvx_OpticalFlowPyrLK_workspace_size(level-0 width, height, capacity) bytes, device,
 * 16-byte aligned) must be zero-filled before the first call and then kept: like the reference
 * node's local data it carries state from call to call that the reference's results depend on. */
size_t mvx_OpticalFlowPyrLK_workspace_size(uint32_t width0, uint32_t height0, uint32_t capacity);
int mvx_OpticalFlowPyrLK_XY_XY(mvx_stream_t stream, void *pNewKeypoints, uint32_t *pNewCount,
                               const mvx_image_level_t *oldPyramid, const mvx_image_level_t *newPyramid, uint32_t levels,
                               float pyramidScale, const void *pOldKeypoints, const uint32_t *pOldCount,
                               const void *pEstimates, uint32_t capacity, int32_t termination, float epsilon,
                               uint32_t numIterations, int32_t useInitialEstimate, uint32_t windowDimension, void *workspace,
                               size_t workspaceSize);

/* ------------------------------------------------------------------ fused vision graph ----- */
/* The RGB vision graph (vx nodes: ColorConvert RGB->IYUV, ChannelExtract Y, ScaleImage half bilinear,
 * GaussianPyramid HALF, CannyEdgeDetector 3/L1, HarrisCorners 3/3) over a batch of equally sized
 * frames in a handful of launches.  It has no counterpart among the reference's HIP launchers: the
 * reference runs one kernel per node and the Canny trace and Harris pick on the CPU
 * (ago_util_hip.cpp:740-744 leaves supernode launches unimplemented).  Outputs are bit-exact with the
 * per-node entry points above.  Applies when mvx_VisionGraph_supported() returns 1 (width >= 16 and
 * a multiple of 4, even height, scaled image exactly half size, valid pyramid steps). */
typedef struct mvx_vision_params {
    uint32_t width, height;              /* RGB input size */
    uint32_t scaledWidth, scaledHeight;  /* ScaleImage output (width/2, height/2) */
    uint32_t levels;                     /* pyramid levels (1..8), level 0 = luma */
    uint32_t cannyLower, cannyUpper;     /* hysteresis thresholds */
    float harrisK, harrisThreshold, harrisMinDistance;
    uint32_t cornerCapacity;             /* vx_keypoint_t capacity of each frame's corner array */
} mvx_vision_params_t;

typedef struct mvx_vision_frame {
    const uint8_t *rgb; uint32_t rgbStride;
    uint8_t *pyramid[8]; uint32_t pyramidStride[8];
    uint8_t *scaled; uint32_t scaledStride;
    uint8_t *edges; uint32_t edgesStride;    /* pixels outside [1,w-2]x[1,h-2] are not written */
    void *corners;                           /* vx_keypoint_t[cornerCapacity] */
    uint32_t *numItems;                      /* device: min(count, capacity) */
    uint64_t *numCorners;                    /* device: vx_size corner count (uncapped) */
} mvx_vision_frame_t;

int mvx_VisionGraph_supported(const mvx_vision_params_t *params);
size_t mvx_VisionGraph_workspace_size(const mvx_vision_params_t *params, uint32_t maxFrames);
/* once per set of frames: validates the planes and writes the frame table into the workspace (device
 * memory, 16-byte aligned); synchronous with respect to the host. */
int mvx_VisionGraph_prepare(mvx_stream_t stream, const mvx_vision_params_t *params, const mvx_vision_frame_t *frames,
                            uint32_t nframes, void *workspace, size_t workspaceSize);
/* every batch: enqueues the graph for the prepared frames; asynchronous, capturable in a hipGraph.
 * The fused tile pass and then the hysteresis stages run on the caller's stream, the pyramid levels and
 * the Harris pick beside the hysteresis on a library-owned stream of the current device (default
 * priority); the caller's stream waits for it before the call's work counts as done. */
int mvx_VisionGraph_run(mvx_stream_t stream, const mvx_vision_params_t *params, uint32_t nframes, void *workspace,
                        size_t workspaceSize);
/* A lane: the two default-priority side streams and their fork / join events on which
 * mvx_VisionGraph_run_lane runs the stages after each chunk's tile pass.  mvx_VisionGraph_run uses one
 * lane per device shared by all its callers; a caller that runs several graphs concurrently gives each
 * its own lane.  Create and destroy outside stream capture. */
typedef struct mvx_vision_lane mvx_vision_lane_t;
int mvx_VisionGraph_lane_create(mvx_vision_lane_t **lane);
int mvx_VisionGraph_lane_destroy(mvx_vision_lane_t *lane);
/* mvx_VisionGraph_run with the side stages on `lane` (NULL: every stage on `stream`, all frames at once) */
int mvx_VisionGraph_run_lane(mvx_stream_t stream, mvx_vision_lane_t *lane, const mvx_vision_params_t *params,
                             uint32_t nframes, void *workspace, size_t workspaceSize);
/* the same in chunks of framesPerChunk frames (0: the default, all frames in one chunk): each chunk's
 * tile pass on the caller's stream, then its later stages on the lane's streams while the caller's
 * stream runs the next chunk's tile pass */
int mvx_VisionGraph_run_pipelined(mvx_stream_t stream, mvx_vision_lane_t *lane, const mvx_vision_params_t *params,
                                  uint32_t nframes, void *workspace, size_t workspaceSize, uint32_t framesPerChunk);
/* the same, one stage at a time (0 fused tile pass, 1 hysteresis border union, 2 hysteresis resolve,
 * 3 pyramid levels 2.., 4 Harris sort and pick).  Stage 0 comes first and 2 after 1; stages 3 and 4
 * need only stage 0 and may run on another stream concurrently with 1 and 2.  Frames are independent:
 * the stages of one frame may overlap another frame's.  The next batch's stage 0 of a frame must follow
 * every stage of that frame in the previous batch. */
#define MVX_VISION_STAGES 5
int mvx_VisionGraph_run_stage(mvx_stream_t stream, const mvx_vision_params_t *params, uint32_t nframes,
                              void *workspace, size_t workspaceSize, int stage);
/* one stage for the prepared frames firstFrame .. firstFrame + nframes - 1 */
int mvx_VisionGraph_run_stage_range(mvx_stream_t stream, const mvx_vision_params_t *params, uint32_t firstFrame,
                                    uint32_t nframes, void *workspace, size_t workspaceSize, int stage);


/* ------------------------------------------------------------------ batched chain ---------- */
/* Gaussian3x3 -> Sobel3x3 GXY -> Magnitude + Phase, and ScaleImage (bilinear) of the input to half
 * size, for a batch of U8 images of uneven sizes in one persistent launch.  Per image the outputs
 * equal those of the node sequence (the blurred image virtual):
 *   mvx_Gaussian_U8_U8_3x3 -> mvx_Sobel_S16S16_U8_3x3_GXY -> mvx_Magnitude_S16_S16S16, mvx_Phase_U8_S16S16,
 *   mvx_ScaleImage_U8_U8_Bilinear(width/2, height/2) of the input,
 * i.e. the reference's HafCpu_Gaussian_U8_U8_3x3 (ago_haf_cpu_filter.cpp:1681-1861),
 * HafCpu_Sobel_S16S16_U8_3x3_GXY (:2237), HafCpu_Magnitude_S16_S16S16 (ago_haf_cpu_arithmetic.cpp:4989),
 * HafCpu_Phase_U8_S16S16 (:9592) and HafCpu_ScaleImage_U8_U8_Bilinear (ago_haf_cpu_geometric.cpp:1977)
 * run one node at a time; the reference has no batched launch (its executor launches one kernel per
 * node and frame, ago_util_hip.cpp:686-738).  Rows 0 and height-1 of gx/gy are not written.
 * gx and gy both NULL: the gradients are virtual and the image takes the merged kernel
 * mvx_SobelMagnitudePhase_S16U8_U8_3x3 instead of the three gradient nodes (the reference's merge rule,
 * ago_drama_remove.cpp:155-163): magnitude columns from width & ~15 on truncate, and rows 0 and height-1
 * of magnitude and phase are not written.
 * Widths and heights are even (2..65536), h*srcStride < 2^31; half may be NULL. */
typedef struct mvx_chain_image {
    const uint8_t *src; uint32_t width, height, srcStride;
    int16_t *gx; uint32_t gxStride;
    int16_t *gy; uint32_t gyStride;
    int16_t *mag; uint32_t magStride;
    uint8_t *phase; uint32_t phaseStride;
    uint8_t *half; uint32_t halfStride;          /* (width/2) x (height/2) U8, or NULL */
} mvx_chain_image_t;

/* tile shapes (width x height): 16x16, 32x16, 32x32, 64x16, 64x32 */
#define MVX_CHAIN_TILE_COUNT 5
/* workgroups per CU tried by the autotuner: 1, 2, 4 and the occupancy limit */
#define MVX_CHAIN_OCCUPANCY_STEPS 4

typedef struct mvx_chain_plan {
    void *table;                                 /* device plan (in the caller's workspace) */
    uint32_t count, numTiles, tile, workgroups;  /* images, tiles, tile shape index, persistent grid */
    uint32_t tileWidth, tileHeight;
} mvx_chain_plan_t;

size_t mvx_ChainBatch_workspace_size(uint32_t count);
/* Validates the images and writes the device plan into `workspace` (device memory, 16-byte
 * aligned); synchronous with respect to the host.  workgroupsPerCU 0 = the occupancy limit. */
int mvx_ChainBatch_prepare(const mvx_chain_image_t *images, uint32_t count, uint32_t tile, uint32_t workgroupsPerCU,
                           void *workspace, size_t workspaceSize, mvx_chain_plan_t *plan);
/* One persistent launch over every tile of the plan; asynchronous, capturable. */
int mvx_ChainBatch_run(mvx_stream_t stream, const mvx_chain_plan_t *plan);
/* The same plan from the amd_rpp batchPD layout (amd_rpp/source/image/<name>batchPD.cpp, e.g.
 * AbsoluteDifferencebatchPD.cpp:52-61, and the rppi_*_batchPD_gpu calls at :146-161): nbatch images
 * stacked in one tall plane per role, image i in rows [i*maxHeight, i*maxHeight + heights[i]) and
 * columns [0, widths[i]) of the plane, with the plane's stride; widths / heights are host arrays (the
 * srcBatch_width / srcBatch_height arrays).  The outputs gx, gy, mag, phase use the same rows per
 * image; half (optional) uses halfMaxHeight rows per image.  Neighbour reads follow the linear memory
 * model with the tall plane's stride (column 0 and w-1 see the plane's bytes beside the image).  No
 * reference counterpart: it lets batchPD callers (RPP-style plug-ins) reach the batched chain without
 * building descriptors. */
int mvx_ChainBatch_prepare_batchPD(const uint8_t *pSrc, uint32_t srcStride, uint32_t maxHeight, const uint32_t *widths,
                                   const uint32_t *heights, uint32_t nbatch, int16_t *pGx, uint32_t gxStride, int16_t *pGy,
                                   uint32_t gyStride, int16_t *pMag, uint32_t magStride, uint8_t *pPhase, uint32_t phaseStride,
                                   uint8_t *pHalf, uint32_t halfStride, uint32_t halfMaxHeight, uint32_t tile,
                                   uint32_t workgroupsPerCU, void *workspace, size_t workspaceSize, mvx_chain_plan_t *plan);
/* Times every tile shape x workgroups-per-CU step (reps launches each, HIP events on `stream`),
 * prepares the fastest into `workspace` and returns its plan; msTable (optional) receives
 * MVX_CHAIN_TILE_COUNT * MVX_CHAIN_OCCUPANCY_STEPS milliseconds per launch.  Synchronises `stream`;
 * the outputs hold the result of the last timed launch. */
int mvx_ChainBatch_autotune(mvx_stream_t stream, const mvx_chain_image_t *images, uint32_t count, uint32_t reps,
                            void *workspace, size_t workspaceSize, mvx_chain_plan_t *best, float *msTable);
/* Captures one mvx_ChainBatch_run of `plan` into a HIP graph and instantiates it (synchronous). */
int mvx_ChainBatch_graph_create(const mvx_chain_plan_t *plan, void **graphExec);
int mvx_graph_launch(void *graphExec, mvx_stream_t stream);
int mvx_graph_destroy(void *graphExec);


/* ------------------------------------------------------------------ differentiable F32 warp -- */
/* F32 bilinear WarpPerspective / Remap with hand-written adjoints, for gradient-based alignment.
 * The reference has no F32 warp and no adjoint; the sampling convention is that of its U8 warps
 * (HafCpu_WarpPerspective_U8_U8_Bilinear_Constant, ago_haf_cpu_geometric.cpp:1579; Remap :249):
 * destination (x, y) samples the source at (u, v) = ((h0 x + h1 y + h2)/z, (h3 x + h4 y + h5)/z),
 * z = h6 x + h7 y + h8, bilinearly; source pixels outside the image read 0; z <= 0 gives 0.
 * `count` images of one size per call, `...BatchStride` bytes apart; strides in bytes, 4-byte
 * aligned; H is device double[count][9] (the kernels compute in F32 from h0 - h8, h4 - h8 and the
 * other entries, the differences formed in F64); gradients w.r.t. the source are ACCUMULATED into gradSrc
 * (may be NULL), dL/dH is written as double[count][9] (d/dh8 included; callers fixing h8 = 1 use
 * the first 8). */
int mvx_WarpPerspective_F32_Bilinear(mvx_stream_t stream, uint32_t count, uint32_t dstWidth, uint32_t dstHeight, float *dst,
                                     uint32_t dstStride, size_t dstBatchStride, uint32_t srcWidth, uint32_t srcHeight,
                                     const float *src, uint32_t srcStride, size_t srcBatchStride, const double *H);
/* device workspace of the adjoint and L2 entry points (per-workgroup partial sums) */
size_t mvx_warp_perspective_f32_workspace_size(uint32_t count, uint32_t dstWidth, uint32_t dstHeight);
/* given dL/d(out), dL/d(src) (scatter-add) and dL/dH (deterministic two-stage sum) */
int mvx_warp_perspective_f32_adjoint(mvx_stream_t stream, uint32_t count, uint32_t dstWidth, uint32_t dstHeight,
                                     const float *gradOut, uint32_t gStride, size_t gBatchStride, uint32_t srcWidth,
                                     uint32_t srcHeight, const float *src, uint32_t srcStride, size_t srcBatchStride,
                                     const double *H, float *gradSrc, uint32_t gsStride, size_t gsBatchStride, double *gradH,
                                     void *workspace, size_t workspaceSize);
/* fused forward + L2 loss + adjoint: loss[i] = weight * sum (warp(src_i, H_i) - target_i)^2 and its
 * gradients; with accumulate != 0 loss and gradH are added to instead of overwritten */
int mvx_warp_perspective_f32_l2(mvx_stream_t stream, uint32_t count, uint32_t dstWidth, uint32_t dstHeight, const float *target,
                                uint32_t tStride, size_t tBatchStride, uint32_t srcWidth, uint32_t srcHeight, const float *src,
                                uint32_t srcStride, size_t srcBatchStride, const double *H, float weight, double *loss,
                                double *gradH, float *gradSrc, uint32_t gsStride, size_t gsBatchStride, int accumulate,
                                void *workspace, size_t workspaceSize);
/* map: float (u, v) per destination pixel (mapStride bytes per row) */
int mvx_Remap_F32_Bilinear(mvx_stream_t stream, uint32_t count, uint32_t dstWidth, uint32_t dstHeight, float *dst,
                           uint32_t dstStride, size_t dstBatchStride, uint32_t srcWidth, uint32_t srcHeight, const float *src,
                           uint32_t srcStride, size_t srcBatchStride, const float *map, uint32_t mapStride, size_t mapBatchStride);
/* dL/d(src) (accumulated, may be NULL) and dL/d(map) (written, may be NULL) */
int mvx_remap_f32_adjoint(mvx_stream_t stream, uint32_t count, uint32_t dstWidth, uint32_t dstHeight, const float *gradOut,
                          uint32_t gStride, size_t gBatchStride, uint32_t srcWidth, uint32_t srcHeight, const float *src,
                          uint32_t srcStride, size_t srcBatchStride, const float *map, uint32_t mapStride, size_t mapBatchStride,
                          float *gradSrc, uint32_t gsStride, size_t gsBatchStride, float *gradMap, uint32_t gmStride,
                          size_t gmBatchStride);
/* F32 half-scale Gaussian pyramid level: out(x, y) = sum w_i w_j in(clamp(2x+i-2), clamp(2y+j-2)) / 256,
 * w = [1 4 6 4 1], dst = ceil(src / 2); and its adjoint (accumulated into gradSrc) */
int mvx_PyramidHalf_F32(mvx_stream_t stream, uint32_t count, uint32_t dstWidth, uint32_t dstHeight, float *dst, uint32_t dstStride,
                        size_t dstBatchStride, uint32_t srcWidth, uint32_t srcHeight, const float *src, uint32_t srcStride,
                        size_t srcBatchStride);
int mvx_pyramid_half_f32_adjoint(mvx_stream_t stream, uint32_t count, uint32_t srcWidth, uint32_t srcHeight, float *gradSrc,
                                 uint32_t gsStride, size_t gsBatchStride, uint32_t dstWidth, uint32_t dstHeight,
                                 const float *gradDst, uint32_t gStride, size_t gBatchStride);

#ifdef __cplusplus
}
#endif
#endif
