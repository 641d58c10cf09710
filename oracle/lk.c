/*
 * lk.c -- CPU restatement of the reference's pyramidal Lucas-Kanade tracker
 * HafCpu_OpticalFlowPyrLK_XY_XY_Generic (ago_haf_cpu_opticalflow.cpp:153-546) with its Scharr pass
 * ComputeSharr (:59-150) and helpers (:24-58), as called by agoKernel_OpticalFlowPyrLK_XY_XY
 * (ago_kernel_api.cpp:20587-20650) (TEST INFRASTRUCTURE: the checker, never the product).
 *
 * The reference keeps three pieces of state in the node's zero-initialised local data across
 * pyramid levels AND across calls, and reads parts of them before writing them:
 *   - the Scharr row scratch B (u16; trow0 = B, trow1 = B + ceil8(W+2) for a level of width W):
 *     the horizontal pass at column 0 reads trow0[-1] (never written: 0) and trow1[-1], and at the
 *     last processed column trow0[W8] and trow1[W8] (W8 = W & ~7); those cells hold what an earlier
 *     level (or call) left there;
 *   - the derivative image D (int16 Ix, Iy per pixel, row stride = level-0 width): each level writes
 *     rows 1..H-2, columns 0..W8-1 only; row 0 and columns W8.. keep older values;
 *   - the per-point float positions between levels.
 * oracle_lk_state holds the same state so that a sequence of calls reproduces the reference's.
 *
 * Float sums follow the reference's SSE order: per-lane partial sums (4 lanes over groups of 4 or
 * 8 window columns, rows in order), scalar columns accumulated separately, lanes combined at the
 * end exactly as the reference writes it; b1/b2 accumulate in double.
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define W_BITS 14
#define FLT_SCALE (1.0f / (1 << 20))
#define LK_EPS_ITER_CAP 10000  /* EPSILON-only termination has no iteration bound in the reference */

static inline int descale(int x, int n) { return (x + (1 << (n - 1))) >> n; }

size_t oracle_lk_state_bytes(int w0, int h0, int capacity)
{
    return (size_t)(2 * w0 + 64) * 2 + (size_t)w0 * h0 * 4 + (size_t)capacity * 8;
}

/* B at offset 8 u16 (trow0 = ALIGN16(tempBuf + 1) with a 32-byte aligned tempBuf), D, next */
typedef struct { uint16_t *B; int16_t *D; float *next; int w0, h0; } lk_state;

static lk_state state_of(void *mem, int w0, int h0)
{
    lk_state s;
    uint16_t *base = (uint16_t *)mem;
    s.B = base + 8;
    s.D = (int16_t *)(base + (2 * w0 + 64));
    s.next = (float *)((uint8_t *)s.D + (size_t)w0 * h0 * 4);
    s.w0 = w0;
    s.h0 = h0;
    return s;
}

/* ComputeSharr on one level: D rows 1..H-2, columns 0..W8-1, through the B scratch */
static void scharr(lk_state *st, int W, int H, const uint8_t *src, int stride)
{
    const int W8 = W & ~7, T1 = (W + 2 + 7) & ~7;
    uint16_t *t0 = st->B, *t1 = st->B + T1;
    for (int y = 1; y < H - 1; y++) {
        const uint8_t *r0 = src + (long)(y - 1) * stride, *r1 = src + (long)y * stride, *r2 = src + (long)(y + 1) * stride;
        for (int x = 0; x < W8; x++) {
            t0[x] = (uint16_t)((r0[x] + r2[x]) * 3 + r1[x] * 10);
            t1[x] = (uint16_t)(r2[x] - r0[x]);
        }
        int16_t *d = st->D + (long)y * st->w0 * 2;
        for (int x = 0; x < W8; x++) {
            const int16_t gx = (int16_t)(t0[x + 1] - t0[x - 1]);
            const int16_t gy = (int16_t)(((int16_t)t1[x - 1] + (int16_t)t1[x + 1]) * 3 + (int16_t)t1[x] * 10);
            d[2 * x] = gx;
            d[2 * x + 1] = gy;
        }
    }
}

static float lane_sum4(const float q[4]) { return ((q[0] + q[1]) + q[2]) + q[3]; }

int oracle_lk_run(void *state_mem, int levels, float pyr_scale, const uint8_t *const *old_pyr, const uint8_t *const *new_pyr,
                  const int *ws, const int *hs, const int *strides, int npts, const oracle_keypoint_t *old_kp,
                  const oracle_keypoint_t *est_kp, oracle_keypoint_t *new_kp, int termination, float epsilon,
                  int num_iterations, int use_initial_estimate, int winsz)
{
    /* VX_TERM_CRITERIA_{ITERATIONS,EPSILON,BOTH} = 0xF000, 0xF001, 0xF002 */
    const int TERM_EPS = 0xF001, TERM_BOTH = 0xF002;
    lk_state st = state_of(state_mem, ws[0], hs[0]);
    const size_t halfWin = (size_t)(winsz >> 1);
    const int dStride = ws[0] * 2;   /* elements */
    int16_t Imat[256], dIMat[512];
    for (int level = levels - 1; level >= 0; level--) {
        const int W = ws[level], H = hs[level];
        const unsigned dWidth = (unsigned)W - 2, dHeight = (unsigned)H - 2;
        const unsigned JWidth = (unsigned)W, JHeight = (unsigned)H;
        const int IStride = strides[level], JStride = strides[level];
        const uint8_t *SrcBase = old_pyr[level], *JBase = new_pyr[level];
        scharr(&st, W, H, SrcBase, IStride);
        const float ptScale = (float)pow((double)pyr_scale, (double)level);
        for (int pt = 0; pt < npts; pt++) {
            oracle_keypoint_t *nk = &new_kp[pt];
            const oracle_keypoint_t *ok = &old_kp[pt];
            if (!ok->tracking_status) {
                nk->x = ok->x; nk->y = ok->y; nk->strength = ok->strength; nk->tracking_status = ok->tracking_status;
                nk->scale = ok->scale; nk->error = ok->error;
                continue;
            }
            float prevx = (float)ok->x * ptScale, prevy = (float)ok->y * ptScale;
            float nextx, nexty;
            if (level == levels - 1) {
                if (use_initial_estimate) {
                    nextx = (float)est_kp[pt].x * ptScale;
                    nexty = (float)est_kp[pt].y * ptScale;
                    nk->strength = est_kp[pt].strength;
                    nk->tracking_status = est_kp[pt].tracking_status;
                    nk->error = est_kp[pt].error;
                } else {
                    nextx = prevx;
                    nexty = prevy;
                    nk->tracking_status = ok->tracking_status;
                    nk->strength = ok->strength;
                }
                st.next[2 * pt] = nextx;
                st.next[2 * pt + 1] = nexty;
            } else {
                st.next[2 * pt] *= 2.0f;
                st.next[2 * pt + 1] *= 2.0f;
                nextx = st.next[2 * pt];
                nexty = st.next[2 * pt + 1];
            }
            if (!nk->tracking_status) continue;
            prevx = prevx - (float)halfWin;
            prevy = prevy - (float)halfWin;
            nextx = nextx - (float)halfWin;
            nexty = nexty - (float)halfWin;
            const int ipx = (int)floorf(prevx), ipy = (int)floorf(prevy);
            int bound = (ipx < 0) | (ipx >= (int)(dWidth - (unsigned)winsz)) | (ipy < 0) | (ipy >= (int)(dHeight - (unsigned)winsz));
            if (bound) {
                if (!level) {
                    nk->x = (int32_t)nextx; nk->y = (int32_t)nexty; nk->tracking_status = 0; nk->error = 0;
                }
                continue;
            }
            float a = prevx - (float)ipx, b = prevy - (float)ipy;
            float A11 = 0, A12 = 0, A22 = 0;
            if (a == 0.0f && b == 0.0f) {
                for (int y = 0; y < winsz; y++) {
                    const uint8_t *src = SrcBase + (long)(y + ipy) * IStride + ipx;
                    const int16_t *dsrc = st.D + (long)(y + ipy) * dStride + ipx * 2;
                    int x = 0;
                    for (; x < winsz - 4; x += 4) {
                        int s11 = 0, s22 = 0, s12 = 0;
                        for (int k = 0; k < 4; k++) {
                            const int ix = dsrc[2 * (x + k)], iy = dsrc[2 * (x + k) + 1];
                            Imat[y * winsz + x + k] = (int16_t)(src[x + k] << 5);
                            dIMat[(y * winsz + x + k) * 2] = (int16_t)ix;
                            dIMat[(y * winsz + x + k) * 2 + 1] = (int16_t)iy;
                            s11 += ix * ix; s22 += iy * iy; s12 += ix * iy;
                        }
                        A11 += (float)s11;
                        A22 += (float)s22;
                        A12 += (float)s12;
                    }
                    for (; x < winsz; x++) {
                        const int ix = dsrc[2 * x], iy = dsrc[2 * x + 1];
                        Imat[y * winsz + x] = (int16_t)(src[x] << 5);
                        dIMat[(y * winsz + x) * 2] = (int16_t)ix;
                        dIMat[(y * winsz + x) * 2 + 1] = (int16_t)iy;
                        A11 += (float)(ix * ix);
                        A12 += (float)(ix * iy);
                        A22 += (float)(iy * iy);
                    }
                }
                A11 *= FLT_SCALE;
                A12 *= FLT_SCALE;
                A22 *= FLT_SCALE;
            } else {
                const int iw00 = (int)((double)(((1.f - a) * (1.f - b)) * (float)(1 << W_BITS)) + 0.5);
                const int iw01 = (int)((double)((a * (1.f - b)) * (float)(1 << W_BITS)) + 0.5);
                const int iw10 = (int)((double)(((1.f - a) * b) * (float)(1 << W_BITS)) + 0.5);
                const int iw11 = (1 << W_BITS) - iw00 - iw01 - iw10;
                float q11[4] = {0, 0, 0, 0}, q12[4] = {0, 0, 0, 0}, q22[4] = {0, 0, 0, 0};
                for (int y = 0; y < winsz; y++) {
                    const uint8_t *src = SrcBase + (long)(y + ipy) * IStride + ipx;
                    const int16_t *dsrc = st.D + (long)(y + ipy) * dStride + ipx * 2;
                    int x = 0;
                    for (; x <= winsz - 4; x += 4) {
                        for (int k = 0; k < 4; k++) {
                            const int xx = x + k;
                            const int iv = (src[xx] * iw00 + src[xx + 1] * iw01 + src[xx + IStride] * iw10 + src[xx + IStride + 1] * iw11 +
                                            (1 << (W_BITS - 5 - 1))) >> (W_BITS - 5);
                            const int ix = (dsrc[2 * xx] * iw00 + dsrc[2 * xx + 2] * iw01 + dsrc[2 * xx + dStride] * iw10 +
                                            dsrc[2 * xx + dStride + 2] * iw11 + (1 << (W_BITS - 1))) >> W_BITS;
                            const int iy = (dsrc[2 * xx + 1] * iw00 + dsrc[2 * xx + 3] * iw01 + dsrc[2 * xx + dStride + 1] * iw10 +
                                            dsrc[2 * xx + dStride + 3] * iw11 + (1 << (W_BITS - 1))) >> W_BITS;
                            const int16_t I16 = (int16_t)(iv > 32767 ? 32767 : iv < -32768 ? -32768 : iv);
                            const int16_t X16 = (int16_t)(ix > 32767 ? 32767 : ix < -32768 ? -32768 : ix);
                            const int16_t Y16 = (int16_t)(iy > 32767 ? 32767 : iy < -32768 ? -32768 : iy);
                            Imat[y * winsz + xx] = I16;
                            dIMat[(y * winsz + xx) * 2] = X16;
                            dIMat[(y * winsz + xx) * 2 + 1] = Y16;
                            const float fx = (float)X16, fy = (float)Y16;
                            q22[k] = q22[k] + fy * fy;
                            q12[k] = q12[k] + fx * fy;
                            q11[k] = q11[k] + fx * fx;
                        }
                    }
                    for (; x < winsz; x++) {
                        const int iv = descale(src[x] * iw00 + src[x + 1] * iw01 + src[x + IStride] * iw10 + src[x + IStride + 1] * iw11,
                                               W_BITS - 5);
                        const int ix = descale(dsrc[2 * x] * iw00 + dsrc[2 * x + 2] * iw01 + dsrc[2 * x + dStride] * iw10 +
                                               dsrc[2 * x + dStride + 2] * iw11, W_BITS);
                        const int iy = descale(dsrc[2 * x + 1] * iw00 + dsrc[2 * x + 3] * iw01 + dsrc[2 * x + dStride + 1] * iw10 +
                                               dsrc[2 * x + dStride + 3] * iw11, W_BITS);
                        Imat[y * winsz + x] = (int16_t)iv;
                        dIMat[(y * winsz + x) * 2] = (int16_t)ix;
                        dIMat[(y * winsz + x) * 2 + 1] = (int16_t)iy;
                        A11 += (float)(ix * ix);
                        A12 += (float)(ix * iy);
                        A22 += (float)(iy * iy);
                    }
                }
                if (winsz >= 4) {
                    A11 += lane_sum4(q11);
                    A12 += lane_sum4(q12);
                    A22 += lane_sum4(q22);
                }
                A11 *= FLT_SCALE;
                A12 *= FLT_SCALE;
                A22 *= FLT_SCALE;
            }
            float D = A11 * A22 - A12 * A12;
            const float minEig = (A22 + A11 - sqrtf((A11 - A22) * (A11 - A22) + 4.f * A12 * A12)) / (float)(2 * winsz * winsz);
            if (minEig < 1.0e-04F || D < 1.0e-07F) {
                if (!level) {
                    nk->x = (int32_t)nextx; nk->y = (int32_t)nexty; nk->tracking_status = 0; nk->error = 0;
                }
                continue;
            }
            D = 1.f / D;
            float prevDx = 0.f, prevDy = 0.f, ddx = 0.f, ddy = 0.f;
            unsigned j = 0;
            while ((j < (unsigned)num_iterations || termination == TERM_EPS) && j < LK_EPS_ITER_CAP) {
                const int inx = (int)floorf(nextx), iny = (int)floorf(nexty);
                bound = (inx < 0) | (inx >= (int)(JWidth - (unsigned)winsz)) | (iny < 0) | (iny >= (int)(JHeight - (unsigned)winsz));
                if (bound) {
                    if (!level) { nk->tracking_status = 0; nk->error = 0; }
                    break;
                }
                a = nextx - (float)inx;
                b = nexty - (float)iny;
                const int iw00 = (int)((double)(((1.f - a) * (1.f - b)) * (float)(1 << W_BITS)) + 0.5);
                const int iw01 = (int)((double)((a * (1.f - b)) * (float)(1 << W_BITS)) + 0.5);
                const int iw10 = (int)((double)(((1.f - a) * b) * (float)(1 << W_BITS)) + 0.5);
                const int iw11 = (1 << W_BITS) - iw00 - iw01 - iw10;
                double ib1 = 0, ib2 = 0;
                float qb0[4] = {0, 0, 0, 0}, qb1[4] = {0, 0, 0, 0};
                for (int y = 0; y < winsz; y++) {
                    const uint8_t *Jptr = JBase + (long)(y + iny) * JStride + inx;
                    const int16_t *Iptr = &Imat[y * winsz];
                    const int16_t *dIptr = &dIMat[y * winsz * 2];
                    int x = 0;
                    for (; x <= winsz - 8; x += 8) {
                        int dv[8];
                        for (int k = 0; k < 8; k++) {
                            const int xx = x + k;
                            int t = (Jptr[xx] * iw00 + Jptr[xx + 1] * iw01 + Jptr[xx + JStride] * iw10 + Jptr[xx + JStride + 1] * iw11 +
                                     (1 << (W_BITS - 5 - 1))) >> (W_BITS - 5);
                            t = t > 32767 ? 32767 : t < -32768 ? -32768 : t;
                            int d = t - Iptr[xx];
                            dv[k] = d > 32767 ? 32767 : d < -32768 ? -32768 : d;
                        }
                        /* lanes: qb0 = [Ix*d(x0), Iy*d(x0), Ix*d(x1), Iy*d(x1)], qb1 the same for x2, x3;
                         * then x4, x5 into qb0 and x6, x7 into qb1 */
                        for (int half = 0; half < 2; half++) {
                            const int base = 4 * half;
                            for (int k = 0; k < 2; k++) {
                                const int p = base + k;
                                qb0[2 * k] = qb0[2 * k] + (float)(dIptr[2 * (x + p)] * dv[p]);
                                qb0[2 * k + 1] = qb0[2 * k + 1] + (float)(dIptr[2 * (x + p) + 1] * dv[p]);
                            }
                            for (int k = 0; k < 2; k++) {
                                const int p = base + 2 + k;
                                qb1[2 * k] = qb1[2 * k] + (float)(dIptr[2 * (x + p)] * dv[p]);
                                qb1[2 * k + 1] = qb1[2 * k + 1] + (float)(dIptr[2 * (x + p) + 1] * dv[p]);
                            }
                        }
                    }
                    for (; x < winsz; x++) {
                        int diff = descale(Jptr[x] * iw00 + Jptr[x + 1] * iw01 + Jptr[x + JStride] * iw10 + Jptr[x + JStride + 1] * iw11,
                                           W_BITS - 5);
                        diff -= Iptr[x];
                        ib1 += (double)(float)(diff * dIptr[2 * x]);
                        ib2 += (double)(float)(diff * dIptr[2 * x + 1]);
                    }
                }
                if (winsz >= 8) {
                    float bb[4];
                    for (int k = 0; k < 4; k++) bb[k] = qb0[k] + qb1[k];
                    ib1 += (double)(bb[0] + bb[2]);
                    ib2 += (double)(bb[1] + bb[3]);
                }
                const float b1 = (float)(ib1 * (double)FLT_SCALE), b2 = (float)(ib2 * (double)FLT_SCALE);
                const float dx = (A12 * b2 - A22 * b1) * D;
                const float dy = (A12 * b1 - A11 * b2) * D;
                nextx += dx;
                nexty += dy;
                if ((dx * dx + dy * dy) <= epsilon && (termination == TERM_EPS || termination == TERM_BOTH)) break;
                if (j > 0 && (double)fabsf(dx + prevDx) < 0.01 && (double)fabsf(dy + prevDy) < 0.01) {
                    ddx = dx * 0.5f;
                    ddy = dy * 0.5f;
                    break;
                }
                prevDx = dx;
                prevDy = dy;
                j++;
            }
            if (!level) {
                nk->x = (int32_t)(nextx + (float)halfWin - ddx + 0.5f);
                nk->y = (int32_t)(nexty + (float)halfWin - ddy + 0.5f);
            } else {
                st.next[2 * pt] = nextx + (float)halfWin - ddx;
                st.next[2 * pt + 1] = nexty + (float)halfWin - ddy;
            }
        }
    }
    return 0;
}
