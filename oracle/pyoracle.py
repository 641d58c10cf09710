"""ctypes front-end of the oracle (TEST INFRASTRUCTURE: the checker, never the product).

* ``Oracle``: the C restatement, oracle/liboracle.so (built by ``make -C oracle``).
* ``Reference``: the reference's own CPU engine driven through oracle/_ref/libref_harness.so,
  present only where oracle/_ref was built from the reference sources.

Every function takes and returns packed numpy arrays (row stride = width * bytes per pixel); the
oracle and the reference harness lay planes out with the reference's stride internally.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libref_harness.so")

# vx_keypoint_t
KEYPOINT = np.dtype([("x", np.int32), ("y", np.int32), ("strength", np.float32), ("scale", np.float32),
                     ("orientation", np.float32), ("tracking_status", np.int32), ("error", np.float32)])


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _c(a):
    return np.ascontiguousarray(a)


def build_oracle():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def pyramid_dims(w, h, levels):
    """HALF pyramid level sizes: ceil(0.5 * previous) (ago_util.cpp:1312-1334)"""
    dims = []
    for _ in range(levels):
        dims.append((w, h))
        w, h = (w + 1) // 2, (h + 1) // 2
    return dims


def split_levels(flat, dims):
    out, off = [], 0
    for (w, h) in dims:
        out.append(flat[off:off + w * h].reshape(h, w))
        off += w * h
    return out


def _plane(src, stride=None):
    """a packed U8 image in the reference layout: stride ALIGN16(w) (or the given one), zero padding"""
    h, w = src.shape
    st = (w + 15) & ~15 if stride is None else int(stride)
    out = np.zeros((h, st), np.uint8)
    out[:, :w] = src
    return out, st


# OpenVX enums (VX/vx.h): interpolation types and border modes
INTERP_NEAREST, INTERP_BILINEAR, INTERP_AREA = 0x4000, 0x4001, 0x4002
BORDER_UNDEFINED, BORDER_CONSTANT, BORDER_REPLICATE = 0xC000, 0xC001, 0xC002


def _f32(m):
    return np.ascontiguousarray(np.asarray(m, np.float32).reshape(-1))


# colour formats: the VX fourcc codes and each format's planes as (rows, bytes per row) of a w x h image
def _fourcc(t):
    return ord(t[0]) | ord(t[1]) << 8 | ord(t[2]) << 16 | ord(t[3]) << 24


COLOR_FORMATS = {"RGB": _fourcc("RGB2"), "RGBX": _fourcc("RGBA"), "UYVY": _fourcc("UYVY"), "YUYV": _fourcc("YUYV"),
                 "IYUV": _fourcc("IYUV"), "NV12": _fourcc("NV12"), "NV21": _fourcc("NV21"), "YUV4": _fourcc("YUV4")}
# the conversions the reference's CPU engine runs (ago_drama_divide.cpp's ColorConvert divider)
COLOR_CONVERSIONS = [("RGB", "RGBX"), ("RGB", "IYUV"), ("RGB", "NV12"), ("RGB", "YUV4"),
                     ("RGBX", "RGB"), ("RGBX", "IYUV"), ("RGBX", "NV12"), ("RGBX", "YUV4"),
                     ("UYVY", "RGB"), ("UYVY", "RGBX"), ("UYVY", "IYUV"), ("UYVY", "NV12"),
                     ("YUYV", "RGB"), ("YUYV", "RGBX"), ("YUYV", "IYUV"), ("YUYV", "NV12"),
                     ("IYUV", "RGB"), ("IYUV", "RGBX"), ("IYUV", "NV12"), ("IYUV", "YUV4"),
                     ("NV12", "RGB"), ("NV12", "RGBX"), ("NV12", "IYUV"), ("NV12", "YUV4"),
                     ("NV21", "RGB"), ("NV21", "RGBX"), ("NV21", "IYUV"), ("NV21", "YUV4")]


def color_planes(fmt, w, h):
    """[(rows, bytes per row)] of each plane of a w x h image of `fmt`"""
    return {"RGB": [(h, 3 * w)], "RGBX": [(h, 4 * w)], "UYVY": [(h, 2 * w)], "YUYV": [(h, 2 * w)],
            "IYUV": [(h, w), (h // 2, w // 2), (h // 2, w // 2)], "NV12": [(h, w), (h // 2, w)], "NV21": [(h, w), (h // 2, w)],
            "YUV4": [(h, w), (h, w), (h, w)]}[fmt]


def split_planes(fmt, w, h, packed):
    out, off = [], 0
    for rows, rb in color_planes(fmt, w, h):
        out.append(np.ascontiguousarray(packed[off:off + rows * rb].reshape(rows, rb)))
        off += rows * rb
    return out


def join_planes(planes):
    return np.concatenate([np.ascontiguousarray(p).reshape(-1) for p in planes])


class Oracle:
    def __init__(self):
        if not os.path.exists(ORACLE_SO):
            build_oracle()
        L = self.lib = ctypes.CDLL(ORACLE_SO)
        L.oracle_magnitude_px.restype = ctypes.c_int16
        L.oracle_phase_px.restype = ctypes.c_uint8
        L.oracle_merged_magnitude_px.restype = ctypes.c_int16
        L.oracle_harris_norm.restype = ctypes.c_float
        L.oracle_harris_norm_gb.restype = ctypes.c_float

    # ---- 3x3 filters, Sobel, Magnitude, Phase ----
    def filter3x3(self, which, src):
        h, w = src.shape
        dst = np.zeros((h, w), np.uint8)
        self.lib.oracle_filter3x3(which, w, h, _p(_c(src)), _p(dst))
        return dst

    def sobel(self, src):
        h, w = src.shape
        gx = np.zeros((h, w), np.int16)
        gy = np.zeros((h, w), np.int16)
        self.lib.oracle_sobel_node(w, h, _p(_c(src)), _p(gx), _p(gy))
        return gx, gy

    def magnitude(self, gx, gy):
        h, w = gx.shape
        out = np.zeros((h, w), np.int16)
        self.lib.oracle_magnitude(w, h, _p(out), 2 * w, _p(_c(gx)), 2 * w, _p(_c(gy)), 2 * w)
        return out

    def phase(self, gx, gy):
        h, w = gx.shape
        out = np.zeros((h, w), np.uint8)
        self.lib.oracle_phase(w, h, _p(out), w, _p(_c(gx)), 2 * w, _p(_c(gy)), 2 * w)
        return out

    def chain1(self, src):
        h, w = src.shape
        gx = np.zeros((h, w), np.int16); gy = np.zeros((h, w), np.int16)
        mag = np.zeros((h, w), np.int16); ph = np.zeros((h, w), np.uint8)
        self.lib.oracle_chain1(w, h, _p(_c(src)), _p(gx), _p(gy), _p(mag), _p(ph))
        return gx, gy, mag, ph

    def sobel_merged(self, src, which=0, blur=False):
        """[Gaussian3x3 (virtual) ->] Sobel3x3 (virtual gx, gy) -> Magnitude / Phase as the merge rules run it.
        which: 0 magnitude + phase, 1 magnitude only, 2 phase only.  Returns (mag, phase) with None for the
        output the graph does not have."""
        h, w = src.shape
        mag = np.zeros((h, w), np.int16); ph = np.zeros((h, w), np.uint8)
        self.lib.oracle_sobel_merged_graph(which, 1 if blur else 0, w, h, _p(_c(src)), _p(mag), _p(ph))
        return (mag if which != 2 else None), (ph if which != 1 else None)

    def merged_magnitude_px(self, gx, gy, tail):
        return int(self.lib.oracle_merged_magnitude_px(gx, gy, 1 if tail else 0))

    # ---- colour, scale, pyramid ----
    def color_y(self, rgb):
        h, w = rgb.shape[:2]
        y = np.zeros((h, w), np.uint8)
        self.lib.oracle_color_y_node(w, h, _p(_c(rgb)), _p(y))
        return y

    def scale(self, src, dw, dh):
        sh, sw = src.shape
        d = np.zeros((dh, dw), np.uint8)
        self.lib.oracle_scale_node(sw, sh, _p(_c(src)), dw, dh, _p(d))
        return d

    def scale_mode(self, src, dw, dh, interp, border_mode=BORDER_UNDEFINED, border=0):
        """vxScaleImageNode semantics (ago_drama_divide.cpp:563-616 and the rewrite of
        ago_drama_remove.cpp:1270-1311): same size -> plane copy; otherwise nearest for
        NEAREST_NEIGHBOR, the border form of bilinear for BILINEAR (UNDEFINED / REPLICATE / CONSTANT),
        and for AREA (any border) nearest unless both dimensions shrink, else area.  Unwritten pixels
        stay 0."""
        sh, sw = src.shape
        s = _c(src)
        d = np.zeros((dh, dw), np.uint8)
        if (sw, sh) == (dw, dh):
            d[:] = src
        elif interp == INTERP_NEAREST:
            self.lib.oracle_scale_nearest(dw, dh, _p(d), dw, sw, sh, _p(s), sw)
        elif interp == INTERP_BILINEAR:
            if border_mode == BORDER_UNDEFINED:
                # reads one byte past the row at the right edge: the node wrapper models the reference's
                # zero-padded ALIGN16 rows
                self.lib.oracle_scale_node(sw, sh, _p(s), dw, dh, _p(d))
            else:
                self.lib.oracle_scale_bilinear_border(dw, dh, _p(d), dw, sw, sh, _p(s), sw,
                                                      int(border_mode == BORDER_CONSTANT), int(border))
        elif interp == INTERP_AREA:
            d = self.scale_area(src, dw, dh)
        else:
            raise ValueError(interp)
        return d

    def scale_area(self, src, dw, dh):
        """ScaleImage_U8_U8_Area as a graph runs it: copy at the same size, nearest unless both
        dimensions shrink, the 2x2 average at an exact 2x; other shrinking ratios have no defined result"""
        sh, sw = src.shape
        d = np.zeros((dh, dw), np.uint8)
        if self.lib.oracle_scale_area(dw, dh, _p(d), dw, sw, sh, _p(_c(src)), sw) != 0:
            raise ValueError("a shrinking area scale is defined for an exact 2x (width >= 8) only")
        return d

    def pyramid(self, src, levels):
        h, w = src.shape
        dims = pyramid_dims(w, h, levels)
        flat = np.zeros(sum(a * b for a, b in dims), np.uint8)
        if self.lib.oracle_pyramid_node(w, h, _p(_c(src)), levels, _p(flat)) != 0:
            raise ValueError("pyramid level too small (ScaleGaussianHalf needs src >= 5, dst >= 3)")
        return split_levels(flat, dims)

    def gaussian_half(self, src, dw, dh):
        """one ScaleGaussianHalf_5x5 on reference-layout planes"""
        sh, sw = src.shape
        ss, ds = (sw + 15) & ~15, (dw + 15) & ~15
        s = np.zeros((sh, ss), np.uint8); s[:, :sw] = src
        d = np.zeros((dh, ds), np.uint8)
        if self.lib.oracle_scale_gaussian_half_5x5(dw, dh, _p(d), ds, sw, sh, _p(s), ss) != 0:
            raise ValueError("ScaleGaussianHalf needs src >= 5 and dst >= 3")
        return np.ascontiguousarray(d[:, :dw])

    def color_rgb_yuv(self, y, u, v=None):
        """color_convert IYUV -> RGB (u, v planes) or NV12 -> RGB (u = the interleaved (h/2, w) UV plane)"""
        h, w = y.shape
        nv12 = v is None
        d = np.zeros((h, w, 3), np.uint8)
        vv = _c(v) if v is not None else _c(u)
        self.lib.oracle_color_rgb_yuv(w, h, _p(d), 3 * w, _p(_c(y)), w, _p(_c(u)), u.shape[1], _p(vv), vv.shape[1], int(nv12))
        return d

    def color_iyuv_from_rgb(self, rgb):
        """color_convert RGB -> IYUV, all three planes (ago_haf_cpu_color_convert.cpp:2492-2705)"""
        h, w = rgb.shape[:2]
        y = np.zeros((h, w), np.uint8)
        u = np.zeros((h // 2, w // 2), np.uint8)
        v = np.zeros((h // 2, w // 2), np.uint8)
        self.lib.oracle_color_iyuv_from_rgb(w, h, _p(_c(rgb)), 3 * w, _p(y), w, _p(u), w // 2, _p(v), w // 2)
        return y, u, v

    def color_convert(self, src_fmt, dst_fmt, w, h, planes):
        """color_convert between any two formats the reference converts (COLOR_CONVERSIONS): a list of
        the input's planes (tight rows) -> a list of the output's planes"""
        out = [np.zeros(shape, np.uint8) for shape in color_planes(dst_fmt, w, h)]
        ins = [np.ascontiguousarray(p, dtype=np.uint8) for p in planes]
        st = lambda a: a.shape[1]
        L = self.lib
        yuv = {"IYUV": 0, "NV12": 1, "NV21": 2, "UYVY": 3, "YUYV": 4}
        if dst_fmt in ("RGB", "RGBX") and src_fmt in yuv:
            p = ins + [ins[0]] * (3 - len(ins))
            L.oracle_cc_rgb_from_yuv(yuv[src_fmt], 3 if dst_fmt == "RGB" else 4, w, h, _p(out[0]), st(out[0]), _p(p[0]), st(p[0]),
                                     _p(p[1]), st(p[1]), _p(p[2]), st(p[2]))
        elif src_fmt in ("RGB", "RGBX") and dst_fmt in ("IYUV", "NV12"):
            o = out + [out[1]] * (3 - len(out))
            L.oracle_cc_yuv420_from_rgb(3 if src_fmt == "RGB" else 4, int(dst_fmt == "NV12"), w, h, _p(ins[0]), st(ins[0]), _p(o[0]),
                                        st(o[0]), _p(o[1]), st(o[1]), _p(o[2]), st(o[2]))
        elif src_fmt in ("RGB", "RGBX") and dst_fmt == "YUV4":
            L.oracle_cc_yuv4_from_rgb(3 if src_fmt == "RGB" else 4, w, h, _p(ins[0]), st(ins[0]), _p(out[0]), w, _p(out[1]), w,
                                      _p(out[2]), w)
        elif (src_fmt, dst_fmt) in (("RGB", "RGBX"), ("RGBX", "RGB")):
            L.oracle_cc_rgb_rgbx(int(dst_fmt == "RGBX"), w, h, _p(ins[0]), st(ins[0]), _p(out[0]), st(out[0]))
        else:
            flat = np.zeros(sum(r * b for r, b in color_planes(dst_fmt, w, h)), np.uint8)
            if L.oracle_cc_format(COLOR_FORMATS[src_fmt], COLOR_FORMATS[dst_fmt], w, h, _p(join_planes(ins)), _p(flat)) != 0:
                raise ValueError(f"color_convert {src_fmt} -> {dst_fmt} is not a conversion the reference runs")
            return split_planes(dst_fmt, w, h, flat)
        return out

    def color_luma(self, src_fmt, w, h, planes):
        """the Y plane of a color_convert whose YUV output only channel_extract(Y) reads: ColorConvert_Y_RGB
        (oracle_color_y_rgb) or _Y_RGBX"""
        y = np.zeros((h, w), np.uint8)
        src = np.ascontiguousarray(planes[0], dtype=np.uint8)
        if src_fmt == "RGB":
            self.lib.oracle_color_y_rgb(w, h, _p(y), w, _p(src), src.shape[1])
        else:
            self.lib.oracle_cc_y_rgbx(w, h, _p(src), src.shape[1], _p(y), w)
        return y

    def halfscale(self, src, ksize):
        """halfscale_gaussian (ago_drama_divide.cpp:1471-1488): kernel size 1 -> ScaleImage nearest,
        3 -> ScaleGaussianHalf 3x3, 5 -> ScaleGaussianHalf 5x5, output (w+1)/2 x (h+1)/2 on
        reference-layout planes; unwritten rows stay 0"""
        sh, sw = src.shape
        dw, dh = (sw + 1) // 2, (sh + 1) // 2
        if ksize == 1:
            return self.scale_mode(src, dw, dh, INTERP_NEAREST)
        if ksize == 5:
            return self.gaussian_half(src, dw, dh)
        if ksize != 3:
            raise ValueError(ksize)
        ss, ds = (sw + 15) & ~15, (dw + 15) & ~15
        s = np.zeros((sh, ss), np.uint8); s[:, :sw] = src
        d = np.zeros((dh, ds), np.uint8)
        if self.lib.oracle_scale_gaussian_half_3x3(dw, dh, _p(d), ds, sw, sh, _p(s), ss) != 0:
            raise ValueError("ScaleGaussianHalf 3x3 needs an output at least 2 rows high")
        return np.ascontiguousarray(d[:, :dw])

    # ---- Canny ----
    def canny(self, src, lower, upper):
        h, w = src.shape
        d = np.zeros((h, w), np.uint8)
        self.lib.oracle_canny_node(w, h, _p(_c(src)), lower, upper, _p(d))
        return d

    def canny_ex(self, src, gsize, l2, lower, upper):
        """CannyEdgeDetector for gradient 3/5/7 and L1/L2 with the thresholds the suppression reads"""
        h, w = src.shape
        d = np.zeros((h, w), np.uint8)
        self.lib.oracle_canny_node_ex(w, h, _p(_c(src)), gsize, int(l2), lower, upper, _p(d))
        return d

    def canny_sobel_ex(self, src, gsize, l2, padded=False):
        """the U16 gradient plane (rows of ALIGN16(2w) bytes when padded, so the slot before row y --
        the previous row's last -- is visible)"""
        h, w = src.shape
        s, ss = _plane(src)
        ds = (2 * w + 15) & ~15
        d = np.zeros((h, ds // 2), np.uint16)
        self.lib.oracle_canny_sobel_ex(w, h, _p(d), ds, _p(s), ss, gsize, int(l2))
        return d if padded else np.ascontiguousarray(d[:, :w])

    def canny_sobel(self, src):
        h, w = src.shape
        d = np.zeros((h, w), np.uint16)
        self.lib.oracle_canny_sobel_node(w, h, _p(_c(src)), _p(d))
        return d

    def canny_supp_threshold(self, mag, lower, upper):
        h, w = mag.shape
        d = np.zeros((h, w), np.uint8)
        n = self.lib.oracle_canny_supp_threshold(w, h, _p(d), w, _p(_c(mag)), 2 * w, lower, upper)
        return d, n

    def canny_trace(self, edges):
        h, w = edges.shape
        img = np.ascontiguousarray(edges).copy()
        self.lib.oracle_canny_trace(w, h, _p(img), w)
        return img

    # ---- Harris ----
    def harris(self, src, k, thr, dist, capacity, with_score=False):
        h, w = src.shape
        out = np.zeros(max(capacity, 1), KEYPOINT)
        hvc = np.zeros((h, w), np.float32)
        n = self.lib.oracle_harris_node(w, h, _p(_c(src)), ctypes.c_float(k), ctypes.c_float(thr),
                                        ctypes.c_float(dist), _p(out), capacity, _p(hvc))
        res = (out[:min(n, capacity)], n)
        return res + (hvc,) if with_score else res

    def harris_norm(self, gradient_size=3):
        return self.lib.oracle_harris_norm(gradient_size)

    def fast(self, src, thr, nonmax, capacity):
        """FastCorners: (keypoints, uncapped count)"""
        h, w = src.shape
        s, ss = _plane(src)
        out = np.zeros(max(capacity, 1), KEYPOINT)
        n = self.lib.oracle_fast_corners(w, h, _p(s), ss, ctypes.c_float(thr), int(nonmax), _p(out), capacity)
        return out[:min(n, capacity)], n

    def harris_planes(self, src, g, b, k, thr):
        """(HG3 plane as the raw floats of its ALIGN16(12w)-byte rows, HVC plane (h, w)) of one gradient
        / block size pair"""
        h, w = src.shape
        s, ss = _plane(src)
        hs = (12 * w + 15) & ~15
        hg3 = np.zeros((h, hs // 4), np.float32)
        if g == 3:
            self.lib.oracle_harris_sobel_hg3(w, h, _p(hg3), hs, _p(s), ss)
        else:
            self.lib.oracle_harris_sobel_g(w, h, _p(hg3), hs, _p(s), ss, g)
        vs = (4 * w + 15) & ~15
        vc = np.zeros((h, vs // 4), np.float32)
        norm = ctypes.c_float(self.lib.oracle_harris_norm_gb(g, b))
        if b == 3:
            self.lib.oracle_harris_score(w, h, _p(vc), vs, _p(hg3), hs, ctypes.c_float(k), ctypes.c_float(thr), norm)
        else:
            self.lib.oracle_harris_score_b(w, h, _p(vc), vs, _p(hg3), hs, ctypes.c_float(k), ctypes.c_float(thr), norm, b)
        return hg3, np.ascontiguousarray(vc[:, :w])

    def harris_ex(self, src, g, b, k, thr, dist, capacity, with_score=False):
        """HarrisCorners with gradient size g and block size b (3, 5, 7)"""
        h, w = src.shape
        out = np.zeros(max(capacity, 1), KEYPOINT)
        hvc = np.zeros((h, w), np.float32)
        n = self.lib.oracle_harris_node_ex(w, h, _p(_c(src)), g, b, ctypes.c_float(k), ctypes.c_float(thr),
                                           ctypes.c_float(dist), _p(out), capacity, _p(hvc))
        res = (out[:min(n, capacity)], n)
        return res + (hvc,) if with_score else res

    # ---- warps and remap ----
    def warp_affine(self, src, dw, dh, m, constant=False, border=0, padded=False, src_stride=None):
        """(dst, unpinned); m = the 2x3 VX matrix as stored (6 floats); padded: return the full
        ALIGN16(dw) rows, including the padding pixels the reference's 4-pixel stores write;
        src_stride: a source row pitch other than ALIGN16(width) (zero padding)"""
        sh, sw = src.shape
        s, ss = _plane(src, src_stride)
        ds = (dw + 15) & ~15
        d = np.zeros((dh, ds), np.uint8)
        u = np.zeros((dh, dw), np.uint8)
        self.lib.oracle_warp_affine_bilinear(dw, dh, _p(d), ds, sw, sh, _p(s), ss, _p(_f32(m)), int(constant), border,
                                             _p(u), dw)
        return (d if padded else np.ascontiguousarray(d[:, :dw])), u.astype(bool)

    def warp_perspective(self, src, dw, dh, m, border=0, padded=False, src_stride=None):
        sh, sw = src.shape
        s, ss = _plane(src, src_stride)
        ds = (dw + 15) & ~15
        d = np.zeros((dh, ds), np.uint8)
        u = np.zeros((dh, dw), np.uint8)
        self.lib.oracle_warp_perspective_bilinear(dw, dh, _p(d), ds, sw, sh, _p(s), ss, _p(_f32(m)), border, _p(u), dw)
        return (d if padded else np.ascontiguousarray(d[:, :dw])), u.astype(bool)

    def remap_fixed(self, mapxy, sw, sh):
        """vxSetRemapPoint's u16 (x, y) map from a float (dh, dw, 2) map"""
        dh, dw = mapxy.shape[:2]
        fixed = np.zeros((dh, dw, 2), np.uint16)
        self.lib.oracle_remap_fixed(dw, dh, sw, sh, _p(_f32(mapxy)), _p(fixed))
        return fixed

    def remap(self, src, mapxy, constant=False, border=0, padded=False):
        sh, sw = src.shape
        dh, dw = mapxy.shape[:2]
        fixed = self.remap_fixed(mapxy, sw, sh)
        s, ss = _plane(src)
        ds = (dw + 15) & ~15
        d = np.zeros((dh, ds), np.uint8)
        self.lib.oracle_remap_bilinear(dw, dh, _p(d), ds, sw, sh, _p(s), ss, _p(fixed), int(constant), border)
        return d if padded else np.ascontiguousarray(d[:, :dw])

    def warp_affine_nearest(self, src, dw, dh, m, border=None, valid=None, padded=False):
        """border None: UNDEFINED (0 outside the mask); valid: the input's valid rectangle (start_x,
        start_y, end_x, end_y), the whole image by default"""
        sh, sw = src.shape
        s, ss = _plane(src)
        ds = (dw + 15) & ~15
        d = np.zeros((dh, ds), np.uint8)
        v = np.ascontiguousarray(valid if valid is not None else (0, 0, sw, sh), np.int32)
        self.lib.oracle_warp_affine_nearest(dw, dh, _p(d), ds, sw, sh, _p(s), ss, _p(_f32(m)), border or 0, _p(v))
        return d if padded else np.ascontiguousarray(d[:, :dw])

    def warp_perspective_nearest(self, src, dw, dh, m, border=None, padded=False):
        sh, sw = src.shape
        s, ss = _plane(src)
        ds = (dw + 15) & ~15
        d = np.zeros((dh, ds), np.uint8)
        self.lib.oracle_warp_perspective_nearest(dw, dh, _p(d), ds, sw, sh, _p(s), ss, _p(_f32(m)), border or 0)
        return d if padded else np.ascontiguousarray(d[:, :dw])

    def remap_nearest(self, src, mapxy, constant=False, border=0, padded=False):
        """(dst, unpinned)"""
        sh, sw = src.shape
        dh, dw = mapxy.shape[:2]
        fixed = self.remap_fixed(mapxy, sw, sh)
        s, ss = _plane(src)
        ds = (dw + 15) & ~15
        d = np.zeros((dh, ds), np.uint8)
        u = np.zeros((dh, dw), np.uint8)
        self.lib.oracle_remap_nearest(dw, dh, _p(d), ds, sw, sh, _p(s), ss, _p(fixed), int(constant), border, _p(u), dw)
        return (d if padded else np.ascontiguousarray(d[:, :dw])), u.astype(bool)

    def stitch_merge(self, images, weights):
        """the stitch merge: per channel the in-order float32 sum of (w_k * (1/255)) * p_k, rounded to
        nearest even and saturated (Loom's merge, amd_loomsl/kernels/merge.cpp:223-290, has no CPU form:
        parity unpinned against the reference; pinned by closed forms in the tests)"""
        k255 = np.float32(1.0) / np.float32(255.0)
        acc = np.zeros(images[0].shape, np.float32)
        for img, wt in zip(images, weights):
            wk = (wt.astype(np.float32) * k255)[..., None]
            acc = (acc + (wk * img.astype(np.float32)).astype(np.float32)).astype(np.float32)
        return np.clip(np.rint(acc), 0, 255).astype(np.uint8)

    def stitch_loom(self, stacked, weights, id0, id1, id2, dst=None):
        """Loom's merge with camera-selection maps (oracle_stitch_loom).  stacked: (count*h, w, 3) U8,
        weights: (count*h, w) U8, id0: (h, ceil(w/8)) U8, id1 / id2: (h, ceil(w/8)) U16; dst (h, w, 3) is
        updated in place where the selection writes (a new zero image when None)"""
        count_h, w = weights.shape
        h = id0.shape[0]
        count = count_h // h
        out = np.zeros((h, w, 3), np.uint8) if dst is None else dst
        self.lib.oracle_stitch_loom(w, h, _p(out), 3 * w, count, _p(_c(stacked)), 3 * w, _p(_c(weights)), w, _p(_c(id0)),
                                    id0.shape[1], _p(_c(id1)), 2 * id1.shape[1], _p(_c(id2)), 2 * id2.shape[1])
        return out

    def stitch_loom_remap(self, cams, fixed, weights, id0, id1, id2, dst=None, workers=1):
        """the same stitch with each camera remapped through its 13.3 table on the fly (constant border 0):
        cams: list of (sh, sw, 3) U8; fixed: (count*h, w, 2) U16 stacked tables; rows split over `workers`
        threads"""
        from concurrent.futures import ThreadPoolExecutor
        count = len(cams)
        h = id0.shape[0]
        w = weights.shape[1]
        sh, sw = cams[0].shape[:2]
        out = np.zeros((h, w, 3), np.uint8) if dst is None else dst
        cams = [_c(x) for x in cams]
        ptrs = (ctypes.c_void_p * count)(*[x.ctypes.data for x in cams])
        fixed, weights, id0, id1, id2 = _c(fixed), _c(weights), _c(id0), _c(id1), _c(id2)
        bands = np.linspace(0, h, workers + 1).astype(int)

        def band(i):
            self.lib.oracle_stitch_loom_remap(w, h, int(bands[i]), int(bands[i + 1]), _p(out), 3 * w, count, ptrs, 3 * sw, sw, sh,
                                              _p(fixed), _p(weights), w, _p(id0), id0.shape[1], _p(id1), 2 * id1.shape[1],
                                              _p(id2), 2 * id2.shape[1])
        with ThreadPoolExecutor(workers) as ex:
            list(ex.map(band, range(workers)))
        return out

    # ---- pyramidal Lucas-Kanade ----
    def lk_state(self, w0, h0, capacity):
        """zeroed scratch state of one tracker node (the reference node's local data)"""
        self.lib.oracle_lk_state_bytes.restype = ctypes.c_size_t
        return np.zeros(self.lib.oracle_lk_state_bytes(w0, h0, capacity), np.uint8)

    def lk(self, old_pyr, new_pyr, old_kp, est_kp, termination, epsilon, num_iterations, use_initial_estimate, window,
           state=None, new_kp=None, scale=0.5):
        """one OpticalFlowPyrLK run over lists of pyramid levels (packed U8 arrays); returns the new
        keypoints.  Pass the same `state` to successive calls to reproduce a node run repeatedly."""
        levels = len(old_pyr)
        h0, w0 = old_pyr[0].shape
        n = len(old_kp)
        if state is None:
            state = self.lk_state(w0, h0, max(n, 1))
        planes_o = [_plane(a)[0] for a in old_pyr]
        planes_n = [_plane(a)[0] for a in new_pyr]
        ptr_t = ctypes.c_void_p * levels
        int_t = ctypes.c_int * levels
        po = ptr_t(*[a.ctypes.data for a in planes_o])
        pn = ptr_t(*[a.ctypes.data for a in planes_n])
        ws = int_t(*[a.shape[1] for a in old_pyr])
        hs = int_t(*[a.shape[0] for a in old_pyr])
        st = int_t(*[a.shape[1] for a in planes_o])
        out = np.zeros(n, KEYPOINT) if new_kp is None else np.ascontiguousarray(new_kp).copy()
        old_kp = np.ascontiguousarray(old_kp, KEYPOINT)
        est_kp = np.ascontiguousarray(est_kp, KEYPOINT)
        self.lib.oracle_lk_run(_p(state), levels, ctypes.c_float(scale), po, pn, ws, hs, st, n, _p(old_kp), _p(est_kp), _p(out),
                               termination, ctypes.c_float(epsilon), num_iterations, int(use_initial_estimate), window)
        return out

    # ---- the 4K vision graph ----
    def graph2(self, rgb, dw, dh, levels, lower, upper, k, thr, dist, capacity):
        y = self.color_y(rgb)
        kp, count = self.harris(y, k, thr, dist, capacity)
        return dict(scaled=self.scale(y, dw, dh), pyramid=self.pyramid(y, levels), edges=self.canny(y, lower, upper),
                    corners=kp, count=count)


class Reference:
    """The reference CPU engine; raises FileNotFoundError when oracle/_ref was not built."""

    def __init__(self):
        if not os.path.exists(REF_SO):
            raise FileNotFoundError(REF_SO)
        self.lib = ctypes.CDLL(REF_SO)

    def _ok(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"reference {what} failed: {rc}")

    def filter3x3(self, which, src):
        h, w = src.shape
        dst = np.zeros((h, w), np.uint8)
        self._ok(self.lib.ref_filter3x3(which, w, h, _p(_c(src)), _p(dst)), "filter3x3")
        return dst

    def sobel(self, src):
        h, w = src.shape
        gx = np.zeros((h, w), np.int16)
        gy = np.zeros((h, w), np.int16)
        self._ok(self.lib.ref_sobel3x3(w, h, _p(_c(src)), _p(gx), _p(gy)), "sobel")
        return gx, gy

    def magnitude_phase(self, gx, gy):
        h, w = gx.shape
        mag = np.zeros((h, w), np.int16)
        ph = np.zeros((h, w), np.uint8)
        self._ok(self.lib.ref_magnitude_phase(w, h, _p(_c(gx)), _p(_c(gy)), _p(mag), _p(ph)), "magnitude_phase")
        return mag, ph

    def chain1(self, src, time_iters=0):
        h, w = src.shape
        gx = np.zeros((h, w), np.int16); gy = np.zeros((h, w), np.int16)
        mag = np.zeros((h, w), np.int16); ph = np.zeros((h, w), np.uint8)
        sec = ctypes.c_double(0.0)
        self._ok(self.lib.ref_chain1(w, h, _p(_c(src)), _p(gx), _p(gy), _p(mag), _p(ph),
                                     time_iters, ctypes.byref(sec)), "chain1")
        if time_iters:
            return (gx, gy, mag, ph), sec.value
        return gx, gy, mag, ph

    def chain_half(self, src, time_iters=0):
        """Gaussian (virtual) -> Sobel -> Magnitude + Phase, and ScaleImage to half size, one graph"""
        h, w = src.shape
        gx = np.zeros((h, w), np.int16); gy = np.zeros((h, w), np.int16)
        mag = np.zeros((h, w), np.int16); ph = np.zeros((h, w), np.uint8)
        half = np.zeros((h // 2, w // 2), np.uint8)
        sec = ctypes.c_double(0.0)
        self._ok(self.lib.ref_chain_half(w, h, _p(_c(src)), _p(gx), _p(gy), _p(mag), _p(ph), _p(half),
                                         time_iters, ctypes.byref(sec)), "chain_half")
        out = dict(gx=gx, gy=gy, mag=mag, phase=ph, half=half)
        return (out, sec.value) if time_iters else out

    def sobel_merged(self, src, which=0, blur=False):
        h, w = src.shape
        mag = np.zeros((h, w), np.int16); ph = np.zeros((h, w), np.uint8)
        self._ok(self.lib.ref_sobel_merged(which, 1 if blur else 0, w, h, _p(_c(src)), _p(mag), _p(ph)), "sobel_merged")
        return (mag if which != 2 else None), (ph if which != 1 else None)

    def color_y(self, rgb):
        h, w = rgb.shape[:2]
        y = np.zeros((h, w), np.uint8)
        self._ok(self.lib.ref_color_y(w, h, _p(_c(rgb)), _p(y)), "color_y")
        return y

    def scale(self, src, dw, dh):
        sh, sw = src.shape
        d = np.zeros((dh, dw), np.uint8)
        self._ok(self.lib.ref_scale(sw, sh, _p(_c(src)), dw, dh, _p(d)), "scale")
        return d

    def scale_mode(self, src, dw, dh, interp, border_mode=BORDER_UNDEFINED, border=0):
        sh, sw = src.shape
        d = np.zeros((dh, dw), np.uint8)
        self._ok(self.lib.ref_scale_mode(sw, sh, _p(_c(src)), dw, dh, _p(d), interp, border_mode, border), "scale_mode")
        return d

    def amd_scale_area(self, src, dw, dh):
        sh, sw = src.shape
        d = np.zeros((dh, dw), np.uint8)
        self._ok(self.lib.ref_amd_scale_area(sw, sh, _p(_c(src)), dw, dh, _p(d)), "amd_scale_area")
        return d

    def color_rgb_yuv(self, y, u, v=None):
        h, w = y.shape
        d = np.zeros((h, w, 3), np.uint8)
        vv = _c(v) if v is not None else _c(u)
        self._ok(self.lib.ref_color_rgb_yuv(w, h, _p(_c(y)), _p(_c(u)), _p(vv), int(v is None), _p(d)), "color_rgb_yuv")
        return d

    def color_iyuv_from_rgb(self, rgb):
        h, w = rgb.shape[:2]
        y = np.zeros((h, w), np.uint8)
        u = np.zeros((h // 2, w // 2), np.uint8)
        v = np.zeros((h // 2, w // 2), np.uint8)
        self._ok(self.lib.ref_color_iyuv_from_rgb(w, h, _p(_c(rgb)), _p(y), _p(u), _p(v)), "color_iyuv_from_rgb")
        return y, u, v

    def color_luma(self, src_fmt, yuv_fmt, w, h, planes):
        y = np.zeros((h, w), np.uint8)
        self._ok(self.lib.ref_color_luma(COLOR_FORMATS[src_fmt], COLOR_FORMATS[yuv_fmt], w, h, _p(join_planes(planes)), _p(y)),
                 f"color_luma {src_fmt} -> {yuv_fmt}")
        return y

    def color_convert(self, src_fmt, dst_fmt, w, h, planes):
        out = np.zeros(sum(r * b for r, b in color_planes(dst_fmt, w, h)), np.uint8)
        self._ok(self.lib.ref_color_convert(COLOR_FORMATS[src_fmt], COLOR_FORMATS[dst_fmt], w, h, _p(join_planes(planes)), _p(out)),
                 f"color_convert {src_fmt} -> {dst_fmt}")
        return split_planes(dst_fmt, w, h, out)

    def halfscale(self, src, ksize):
        sh, sw = src.shape
        d = np.zeros(((sh + 1) // 2, (sw + 1) // 2), np.uint8)
        self._ok(self.lib.ref_halfscale(sw, sh, _p(_c(src)), ksize, _p(d)), "halfscale")
        return d

    def pyramid(self, src, levels):
        h, w = src.shape
        dims = pyramid_dims(w, h, levels)
        flat = np.zeros(sum(a * b for a, b in dims), np.uint8)
        self._ok(self.lib.ref_pyramid(w, h, _p(_c(src)), levels, _p(flat)), "pyramid")
        return split_levels(flat, dims)

    def canny(self, src, lower, upper):
        h, w = src.shape
        d = np.zeros((h, w), np.uint8)
        self._ok(self.lib.ref_canny(w, h, _p(_c(src)), lower, upper, _p(d)), "canny")
        return d

    def canny_ex(self, src, gsize, l2, lower, upper, reps=1):
        """(edges, threshold values after the runs)"""
        h, w = src.shape
        d = np.zeros((h, w), np.uint8)
        lo, hi = ctypes.c_int(0), ctypes.c_int(0)
        self._ok(self.lib.ref_canny_ex(w, h, _p(_c(src)), gsize, int(l2), lower, upper, reps, _p(d), ctypes.byref(lo),
                                       ctypes.byref(hi)), "canny_ex")
        return d, (lo.value, hi.value)

    def harris(self, src, k, thr, dist, capacity):
        h, w = src.shape
        out = np.zeros(max(capacity, 1), KEYPOINT)
        n = ctypes.c_int(0)
        cnt = ctypes.c_long(0)
        self._ok(self.lib.ref_harris(w, h, _p(_c(src)), ctypes.c_float(k), ctypes.c_float(thr), ctypes.c_float(dist),
                                     capacity, _p(out), ctypes.byref(n), ctypes.byref(cnt)), "harris")
        return out[:n.value], cnt.value

    def fast(self, src, thr, nonmax, capacity):
        h, w = src.shape
        out = np.zeros(max(capacity, 1), KEYPOINT)
        n = ctypes.c_int(0)
        cnt = ctypes.c_long(0)
        self._ok(self.lib.ref_fast(w, h, _p(_c(src)), ctypes.c_float(thr), int(nonmax), capacity, _p(out), ctypes.byref(n),
                                   ctypes.byref(cnt)), "fast")
        return out[:n.value], cnt.value

    def harris_ex(self, src, g, b, k, thr, dist, capacity):
        h, w = src.shape
        out = np.zeros(max(capacity, 1), KEYPOINT)
        n = ctypes.c_int(0)
        cnt = ctypes.c_long(0)
        self._ok(self.lib.ref_harris_ex(w, h, _p(_c(src)), ctypes.c_float(k), ctypes.c_float(thr), ctypes.c_float(dist), g, b,
                                        capacity, _p(out), ctypes.byref(n), ctypes.byref(cnt)), "harris_ex")
        return out[:n.value], cnt.value

    def graph2(self, rgb, dw, dh, levels, lower, upper, k, thr, dist, capacity, time_iters=0):
        h, w = rgb.shape[:2]
        dims = pyramid_dims(w, h, levels)
        scaled = np.zeros((dh, dw), np.uint8)
        flat = np.zeros(sum(a * b for a, b in dims), np.uint8)
        edges = np.zeros((h, w), np.uint8)
        kp = np.zeros(max(capacity, 1), KEYPOINT)
        n = ctypes.c_int(0)
        cnt = ctypes.c_long(0)
        sec = ctypes.c_double(0)
        self._ok(self.lib.ref_graph2(w, h, _p(_c(rgb)), dw, dh, _p(scaled), levels, _p(flat), lower, upper, _p(edges),
                                     ctypes.c_float(k), ctypes.c_float(thr), ctypes.c_float(dist), capacity, _p(kp),
                                     ctypes.byref(n), ctypes.byref(cnt), time_iters, ctypes.byref(sec)), "graph2")
        out = dict(scaled=scaled, pyramid=split_levels(flat, dims), edges=edges, corners=kp[:n.value], count=cnt.value)
        if time_iters:
            out["sec_per_frame"] = sec.value
        return out

    def warp(self, kind, src, dw, dh, m, constant=False, border=0, time_iters=0):
        sh, sw = src.shape
        d = np.zeros((dh, dw), np.uint8)
        sec = ctypes.c_double(0)
        self._ok(self.lib.ref_warp(kind, sw, sh, _p(_c(src)), dw, dh, _p(_f32(m)), int(constant), border, _p(d), time_iters,
                                   ctypes.byref(sec)), "warp")
        return (d, sec.value) if time_iters else d

    def remap(self, src, mapxy, constant=False, border=0, time_iters=0):
        sh, sw = src.shape
        dh, dw = mapxy.shape[:2]
        d = np.zeros((dh, dw), np.uint8)
        sec = ctypes.c_double(0)
        self._ok(self.lib.ref_remap(sw, sh, _p(_c(src)), dw, dh, _p(_f32(mapxy)), int(constant), border, _p(d), time_iters,
                                    ctypes.byref(sec)), "remap")
        return (d, sec.value) if time_iters else d

    def warp_padded(self, kind, interp, src, dw, dh, m, constant=False, border=0):
        """WarpAffine (kind 0) / WarpPerspective (kind 1) output rows of ALIGN16(dw) bytes, padding included"""
        sh, sw = src.shape
        d = np.zeros((dh, (dw + 15) & ~15), np.uint8)
        self._ok(self.lib.ref_warp_padded(kind, interp, sw, sh, _p(_c(src)), dw, dh, _p(_f32(m)), int(constant), border,
                                          _p(d)), "warp_padded")
        return d

    def remap_padded(self, interp, src, mapxy, constant=False, border=0):
        sh, sw = src.shape
        dh, dw = mapxy.shape[:2]
        d = np.zeros((dh, (dw + 15) & ~15), np.uint8)
        self._ok(self.lib.ref_remap_padded(interp, sw, sh, _p(_c(src)), dw, dh, _p(_f32(mapxy)), int(constant), border,
                                           _p(d)), "remap_padded")
        return d

    def lk(self, old_img, new_img, levels, old_kp, est_kp, termination, epsilon, num_iterations, use_initial_estimate,
           window, reps=1, timed=False):
        h, w = old_img.shape
        n = len(old_kp)
        out = np.zeros(n, KEYPOINT)
        sec = ctypes.c_double(0)
        self._ok(self.lib.ref_lk(w, h, _p(_c(old_img)), _p(_c(new_img)), levels, n, _p(np.ascontiguousarray(old_kp, KEYPOINT)),
                                 _p(np.ascontiguousarray(est_kp, KEYPOINT)), _p(out), termination, ctypes.c_float(epsilon),
                                 num_iterations, int(use_initial_estimate), window, reps, ctypes.byref(sec)), "lk")
        return (out, sec.value) if timed else out
