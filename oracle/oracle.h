/*
 * oracle.h -- CPU restatement of the reference's CPU (ago_haf_cpu*) algorithms for the
 * OpenVX vision-node hot path.
 *
 * THIS IS TEST INFRASTRUCTURE.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only as the checker.  The product path (libmvx*.so)
 * never links or calls it.
 *
 * Memory model shared by every function (it is the reference's, ago_util.cpp:187-200,1107):
 *   - an image plane is `h` rows of `stride` bytes; pixel (x, y) lives at byte y*stride + x*bpp;
 *   - neighbour reads are LINEAR: the left neighbour of (0, y) is the byte before the row, i.e.
 *     the last byte of row y-1's stride, and the right neighbour of (w-1, y) is the first byte
 *     of row y+1 when stride == w*bpp (ago_haf_cpu_filter.cpp:1706-1707);
 *   - reads before byte 0 or at/after byte h*stride see the zero guard bytes of the reference's
 *     calloc'd allocation (AGO_MEMORY_ALLOC_EXTRA_PADDING, ago_internal.h:57);
 *   - 3x3 filters write rows 1..h-2 only; the caller passes zero-filled outputs so that rows 0
 *     and h-1 read back as the reference's untouched calloc'd rows.
 *
 * Parity status of each function is recorded beside its definition ("pinned by ...").
 */
#ifndef MVX_ORACLE_H
#define MVX_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- filters (ago_haf_cpu_filter.cpp) ---- */
void oracle_gaussian3x3(int w, int h, uint8_t *dst, int dst_stride, const uint8_t *src, int src_stride);
void oracle_box3x3(int w, int h, uint8_t *dst, int dst_stride, const uint8_t *src, int src_stride);
void oracle_median3x3(int w, int h, uint8_t *dst, int dst_stride, const uint8_t *src, int src_stride);
/* gx or gy may be NULL; strides in bytes */
void oracle_sobel3x3(int w, int h, int16_t *gx, int gx_stride, int16_t *gy, int gy_stride,
                     const uint8_t *src, int src_stride);

/* merged Sobel -> Magnitude / Phase (virtual gx, gy); mag or phase may be NULL */
void oracle_sobel_magnitude_phase(int w, int h, int16_t *mag, int mag_stride, uint8_t *phase, int phase_stride,
                                  const uint8_t *src, int src_stride);
void oracle_sobel_phase_merged(int w, int h, uint8_t *phase, int phase_stride, const uint8_t *src, int src_stride);
int16_t oracle_merged_magnitude_px(int gx, int gy, int tail);

/* ---- arithmetic (ago_haf_cpu_arithmetic.cpp) ---- */
void oracle_magnitude(int w, int h, int16_t *mag, int mag_stride, const int16_t *gx, int gx_stride,
                      const int16_t *gy, int gy_stride);
void oracle_phase(int w, int h, uint8_t *phase, int phase_stride, const int16_t *gx, int gx_stride,
                  const int16_t *gy, int gy_stride);
/* scalar forms, exported for exhaustive tests */
int16_t oracle_magnitude_px(int gx, int gy);
uint8_t oracle_phase_px(int gx, int gy);

/* ---- colour, scaling, pyramid (geometric.c) ---- */
void oracle_color_y_rgb(int w, int h, uint8_t *dst, int dst_stride, const uint8_t *rgb, int rgb_stride);
void oracle_scale_bilinear(int dw, int dh, uint8_t *dst, int dst_stride, int sw, int sh, const uint8_t *src,
                           int src_stride);
void oracle_color_rgb_yuv(int w, int h, uint8_t *dst, int dst_stride, const uint8_t *y, int y_stride, const uint8_t *u,
                          int u_stride, const uint8_t *v, int v_stride, int nv12);
int oracle_scale_gaussian_half_3x3(int dw, int dh, uint8_t *dst, int dst_stride, int sw, int sh,
                                   const uint8_t *src, int src_stride);
int oracle_scale_gaussian_half_5x5(int dw, int dh, uint8_t *dst, int dst_stride, int sw, int sh,
                                   const uint8_t *src, int src_stride);
void oracle_scale_nearest(int dw, int dh, uint8_t *dst, int dst_stride, int sw, int sh, const uint8_t *src,
                          int src_stride);
int oracle_scale_border_split(int dw, int sw, int xoffs, int xinc);
void oracle_scale_bilinear_border(int dw, int dh, uint8_t *dst, int dst_stride, int sw, int sh, const uint8_t *src,
                                  int src_stride, int constant, int border);
int oracle_scale_area(int dw, int dh, uint8_t *dst, int dst_stride, int sw, int sh, const uint8_t *src,
                      int src_stride);

/* ---- Canny and Harris (vision.c) ---- */
typedef struct { int16_t x, y; float s; } oracle_xys_t;                 /* ago_keypoint_xys_t */
typedef struct { int32_t x, y; float strength, scale, orientation; int32_t tracking_status; float error; }
    oracle_keypoint_t;                                                   /* vx_keypoint_t */
void oracle_canny_sobel_3x3_l1(int w, int h, uint16_t *dst, int dst_stride, const uint8_t *src, int src_stride);
int oracle_canny_supp_threshold(int w, int h, uint8_t *dst, int dst_stride, const uint16_t *mag, int mag_stride,
                                int lower, int upper);
void oracle_canny_trace(int w, int h, uint8_t *img, int stride);
void oracle_canny_sobel_ex(int w, int h, uint16_t *dst, int dst_stride, const uint8_t *src, int src_stride, int gsize,
                           int l2);
void oracle_canny_node_ex(int w, int h, const uint8_t *src, int gsize, int l2, int lower, int upper, uint8_t *dst);
void oracle_harris_sobel_hg3(int w, int h, float *dst, int dst_stride, const uint8_t *src, int src_stride);
void oracle_harris_score(int w, int h, float *dst, int dst_stride, const float *hg3, int hg3_stride,
                         float k, float thr, float norm);
float oracle_harris_norm(int gradient_size);
int oracle_fast_corners(int w, int h, const uint8_t *src, int stride, float threshold, int nonmax,
                        oracle_keypoint_t *out, int capacity);
void oracle_harris_sobel_g(int w, int h, float *dst, int dst_stride, const uint8_t *src, int src_stride, int g);
void oracle_harris_score_b(int w, int h, float *dst, int dst_stride, const float *hg3, int hg3_stride, float k,
                           float thr, float norm, int b);
float oracle_harris_norm_gb(int gradient_size, int block_size);
int oracle_nonmax_xys(int w, int h, const float *img, int stride, oracle_xys_t *list, int capacity);
int oracle_harris_pick(oracle_xys_t *list, int n, float min_distance, int width, int height,
                       oracle_keypoint_t *out, int capacity);

/* ---- warps and remap (warp.c); m is the VX matrix as stored (2x3 affine: 6 floats, 3x3: 9) ---- */
int oracle_warp_affine_needs_border(int dw, int dh, int sw, int sh, const float m[6]);
void oracle_warp_affine_bilinear(int dw, int dh, uint8_t *dst, int dst_stride, int sw, int sh, const uint8_t *src,
                                 int src_stride, const float m[6], int constant, int border, uint8_t *unpinned,
                                 int unpinned_stride);
void oracle_warp_perspective_bilinear(int dw, int dh, uint8_t *dst, int dst_stride, int sw, int sh, const uint8_t *src,
                                      int src_stride, const float m[9], int border, uint8_t *unpinned, int unpinned_stride);
void oracle_remap_fixed(int dw, int dh, int sw, int sh, const float *map, uint16_t *fixed);
void oracle_warp_affine_nearest(int dw, int dh, uint8_t *dst, int dst_stride, int sw, int sh, const uint8_t *src,
                                int src_stride, const float m[6], int border, const int valid[4]);
void oracle_warp_perspective_nearest(int dw, int dh, uint8_t *dst, int dst_stride, int sw, int sh, const uint8_t *src,
                                     int src_stride, const float m[9], int border);
void oracle_remap_nearest(int dw, int dh, uint8_t *dst, int dst_stride, int sw, int sh, const uint8_t *src,
                          int src_stride, const uint16_t *fixed, int constant, int border, uint8_t *unpinned,
                          int unpinned_stride);

void oracle_remap_bilinear(int dw, int dh, uint8_t *dst, int dst_stride, int sw, int sh, const uint8_t *src,
                           int src_stride, const uint16_t *fixed, int constant, int border);

/* ---- stitch: Loom's merge with camera-selection maps (warp.c); camId1/2 strides in bytes ---- */
void oracle_stitch_loom(int w, int h, uint8_t *dst, int dst_stride, int count, const uint8_t *stacked, int stacked_stride,
                        const uint8_t *wt, int wt_stride, const uint8_t *id0, int id0_stride, const uint16_t *id1, int id1_stride,
                        const uint16_t *id2, int id2_stride);
void oracle_stitch_loom_remap(int w, int h, int y0, int y1, uint8_t *dst, int dst_stride, int count, const uint8_t *const *cams,
                              int cam_stride, int sw, int sh, const uint16_t *fixed, const uint8_t *wt, int wt_stride, const uint8_t *id0,
                              int id0_stride, const uint16_t *id1, int id1_stride, const uint16_t *id2, int id2_stride);

/* ---- pyramidal Lucas-Kanade (lk.c); the state block plays the reference node's local data ---- */
size_t oracle_lk_state_bytes(int w0, int h0, int capacity);
int oracle_lk_run(void *state_mem, int levels, float pyr_scale, const uint8_t *const *old_pyr, const uint8_t *const *new_pyr,
                  const int *ws, const int *hs, const int *strides, int npts, const oracle_keypoint_t *old_kp,
                  const oracle_keypoint_t *est_kp, oracle_keypoint_t *new_kp, int termination, float epsilon,
                  int num_iterations, int use_initial_estimate, int winsz);

/* ---- nodes and graphs on reference-layout planes (graphs.c); packed in/out ---- */
void oracle_filter3x3(int which, int w, int h, const uint8_t *src, uint8_t *dst);
void oracle_sobel_node(int w, int h, const uint8_t *src, int16_t *gx, int16_t *gy);
/* [Gaussian3x3 (virtual) ->] Sobel3x3 (virtual gx, gy) -> Magnitude and/or Phase, as the merge
 * rules run it; which: 0 SobelMagnitudePhase, 1 SobelMagnitude, 2 SobelPhase */
void oracle_sobel_merged_graph(int which, int blur, int w, int h, const uint8_t *src, int16_t *mag, uint8_t *phase);
/* Gaussian3x3 (virtual) -> Sobel3x3 -> Magnitude + Phase */
void oracle_chain1(int w, int h, const uint8_t *src, int16_t *gx, int16_t *gy, int16_t *mag,
                   uint8_t *phase);
void oracle_color_y_node(int w, int h, const uint8_t *rgb, uint8_t *y);
void oracle_scale_node(int sw, int sh, const uint8_t *src, int dw, int dh, uint8_t *dst);
void oracle_pyramid_dims(int w, int h, int levels, int *ws, int *hs);
int oracle_pyramid_node(int w, int h, const uint8_t *src, int levels, uint8_t *out);
void oracle_canny_node(int w, int h, const uint8_t *src, int lower, int upper, uint8_t *dst);
void oracle_canny_sobel_node(int w, int h, const uint8_t *src, uint16_t *dst);
int oracle_harris_node_ex(int w, int h, const uint8_t *src, int g, int b, float k, float thr, float min_distance,
                          oracle_keypoint_t *out, int capacity, float *hvc);
int oracle_harris_node(int w, int h, const uint8_t *src, float k, float thr, float min_distance,
                       oracle_keypoint_t *out, int capacity, float *hvc);

void oracle_color_iyuv_from_rgb(int w, int h, const uint8_t *rgb, int rgbStride, uint8_t *y, int yStride, uint8_t *u,
                                int uStride, uint8_t *v, int vStride);
/* colour conversions beyond RGB <-> IYUV (color.c); formats of the YUV inputs */
enum { ORACLE_FMT_IYUV = 0, ORACLE_FMT_NV12 = 1, ORACLE_FMT_NV21 = 2, ORACLE_FMT_UYVY = 3, ORACLE_FMT_YUYV = 4 };
void oracle_cc_rgb_from_yuv(int fmt, int out_bpp, int w, int h, uint8_t *dst, int dst_stride, const uint8_t *p0, int s0,
                            const uint8_t *p1, int s1, const uint8_t *p2, int s2);
void oracle_cc_yuv420_from_rgb(int in_bpp, int nv12, int w, int h, const uint8_t *src, int src_stride, uint8_t *yp, int ys,
                               uint8_t *up, int us, uint8_t *vp, int vs);
void oracle_cc_yuv4_from_rgb(int in_bpp, int w, int h, const uint8_t *src, int src_stride, uint8_t *yp, int ys, uint8_t *up,
                             int us, uint8_t *vp, int vs);
void oracle_cc_rgb_rgbx(int to_rgbx, int w, int h, const uint8_t *src, int src_stride, uint8_t *dst, int dst_stride);
int oracle_cc_format(uint32_t in_fmt, uint32_t out_fmt, int w, int h, const uint8_t *in, uint8_t *out);
void oracle_cc_y_rgbx(int w, int h, const uint8_t *src, int src_stride, uint8_t *dst, int dst_stride);
#ifdef __cplusplus
}

#endif
#endif
