"""Kernel sequences of the vision graphs, launched through the C ABI on one stream.

These mirror what the reference's graph optimizer schedules for the same OpenVX graphs
(divide: ago_drama_divide.cpp; merges: ago_drama_remove.cpp) and are what the parity tests and the
benchmark drive.  Buffers are device planes in the reference layout; workspaces are device tensors.
"""
import ctypes

import numpy as np
import torch

from . import kernels
from .device import Plane, call

KEYPOINT = np.dtype([("x", np.int32), ("y", np.int32), ("strength", np.float32), ("scale", np.float32),
                     ("orientation", np.float32), ("tracking_status", np.int32), ("error", np.float32)])


def pyramid_dims(w, h, levels):
    dims = []
    for _ in range(levels):
        dims.append((w, h))
        w, h = (w + 1) // 2, (h + 1) // 2
    return dims


def _dev_bytes(n):
    return torch.zeros(max(int(n), 1), dtype=torch.uint8, device="cuda")


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


class Canny:
    """CannyEdgeDetector (3x3, L1): CannySobel -> CannySuppThreshold -> CannyEdgeTrace."""

    def __init__(self, w, h):
        self.w, self.h = w, h
        self.mag = Plane(w, h, "U16")
        self.stack = _dev_bytes(4 * w * h)

    def run(self, src, dst, lower, upper):
        w, h = self.w, self.h
        call("CannySobel_U16_U8_3x3_L1NORM", w, h, self.mag, self.mag.stride, src, src.stride)
        call("CannySuppThreshold_U8XY_U16_3x3", w, h, dst, dst.stride, self.mag, self.mag.stride,
             _ptr(self.stack), w * h, lower, upper)
        call("CannyEdgeTrace_U8_U8XY", w, h, dst, dst.stride, _ptr(self.stack), w * h)


class Harris:
    """HarrisCorners (gradient 3, block 3): HG3 -> score -> NMS -> sort and pick."""

    def __init__(self, w, h, capacity):
        K = kernels()
        K.mvx_NonMaxSupp_XY_ANY_3x3_workspace_size.restype = ctypes.c_size_t
        K.mvx_HarrisMergeSortAndPick_XY_XYS_workspace_size.restype = ctypes.c_size_t
        self.w, self.h, self.capacity = w, h, capacity
        self.hg3 = Plane(w, h, "F32x3")
        self.hvc = Plane(w, h, "F32")
        # NMS maxima are never 8-adjacent: at most one per 2x2 block
        self.max_list = max(1, ((w - 1) // 2) * ((h - 1) // 2))
        self.xys = _dev_bytes(8 * self.max_list)
        self.xys_count = torch.zeros(1, dtype=torch.int32, device="cuda")
        self.nms_ws_size = K.mvx_NonMaxSupp_XY_ANY_3x3_workspace_size(w, h)
        self.nms_ws = _dev_bytes(self.nms_ws_size)
        self.pick_ws_size = K.mvx_HarrisMergeSortAndPick_XY_XYS_workspace_size(self.max_list)
        self.pick_ws = _dev_bytes(self.pick_ws_size)
        call("HarrisMergeSortAndPick_workspace_init", self.max_list, _ptr(self.pick_ws), ctypes.c_size_t(self.pick_ws_size))
        self.corners = _dev_bytes(28 * max(capacity, 1))
        self.num_items = torch.zeros(1, dtype=torch.int32, device="cuda")
        self.count = torch.zeros(1, dtype=torch.int64, device="cuda")

    def run(self, src, k, thr, dist, norm):
        w, h = self.w, self.h
        call("HarrisSobel_HG3_U8_3x3", w, h, self.hg3, self.hg3.stride, src, src.stride)
        call("HarrisScore_HVC_HG3_3x3", w, h, self.hvc, self.hvc.stride, self.hg3, self.hg3.stride,
             float(k), float(thr), float(norm))
        call("NonMaxSupp_XY_ANY_3x3", self.max_list, _ptr(self.xys), _ptr(self.xys_count), w, h, self.hvc,
             self.hvc.stride, _ptr(self.nms_ws), ctypes.c_size_t(self.nms_ws_size))
        call("HarrisMergeSortAndPick_XY_XYS", self.capacity, _ptr(self.corners), _ptr(self.num_items),
             _ptr(self.count), _ptr(self.xys), _ptr(self.xys_count), self.max_list, float(dist), w, h,
             _ptr(self.pick_ws), ctypes.c_size_t(self.pick_ws_size))

    def result(self):
        n = int(self.num_items.item())
        raw = self.corners[:28 * n].cpu().numpy()
        return raw.view(KEYPOINT).copy(), int(self.count.item())


def harris_norm(gradient_size=3):
    """(255 * 2^(g-1) * 3)^4 in float, left to right (ago_kernel_api.cpp:16941-16942)"""
    f = np.float32(255.0) * np.float32(1 << (gradient_size - 1)) * np.float32(3)
    return float(np.float32(np.float32(np.float32(f * f) * f) * f))


class VisionGraph4K:
    """RGB -> IYUV (virtual) -> Y (virtual); Y -> ScaleImage bilinear; Y -> GaussianPyramid HALF;
    Y -> Canny (3, L1); Y -> Harris (3, 3).  The reference keeps ColorConvert_Y_RGB only, a channel
    copy for pyramid level 0, ScaleGaussianHalf_5x5 per further level, and the Canny / Harris kernel
    chains (ago_drama_divide.cpp:244-256, 563-616, 900-922, 1618-1701, 1703-1787)."""

    def __init__(self, w, h, dw, dh, levels=5, lower=80, upper=100, k=0.04, thr=0.0005, dist=5.0, capacity=65536):
        self.w, self.h, self.dw, self.dh, self.levels = w, h, dw, dh, levels
        self.lower, self.upper, self.k, self.thr, self.dist = lower, upper, k, thr, dist
        self.rgb = Plane(w, h, "RGB")
        self.y = Plane(w, h, "U8")
        self.scaled = Plane(dw, dh, "U8")
        self.pyr = [Plane(a, b, "U8") for (a, b) in pyramid_dims(w, h, levels)]
        self.edges = Plane(w, h, "U8")
        self.canny = Canny(w, h)
        self.harris = Harris(w, h, capacity)
        self.norm = harris_norm(3)

    def run(self, rgb_plane=None):
        rgb = rgb_plane if rgb_plane is not None else self.rgb
        w, h = self.w, self.h
        call("ColorConvert_Y_RGB", w, h, self.y, self.y.stride, rgb, rgb.stride)
        call("ScaleImage_U8_U8_Bilinear", self.dw, self.dh, self.scaled, self.scaled.stride, w, h, self.y, self.y.stride)
        call("ChannelCopy_U8_U8", w, h, self.pyr[0], self.pyr[0].stride, self.y, self.y.stride)
        for l in range(1, self.levels):
            a, b = self.pyr[l], self.pyr[l - 1]
            call("ScaleGaussianHalf_U8_U8_5x5", a.width, a.height, a, a.stride, b.width, b.height, b, b.stride)
        self.canny.run(self.y, self.edges, self.lower, self.upper)
        self.harris.run(self.y, self.k, self.thr, self.dist, self.norm)

    def results(self):
        kp, count = self.harris.result()
        return dict(scaled=self.scaled.download(), pyramid=[p.download() for p in self.pyr],
                    edges=self.edges.download(), corners=kp, count=count)


class VisionParams(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32), ("scaledWidth", ctypes.c_uint32),
                ("scaledHeight", ctypes.c_uint32), ("levels", ctypes.c_uint32), ("cannyLower", ctypes.c_uint32),
                ("cannyUpper", ctypes.c_uint32), ("harrisK", ctypes.c_float), ("harrisThreshold", ctypes.c_float),
                ("harrisMinDistance", ctypes.c_float), ("cornerCapacity", ctypes.c_uint32)]


class VisionFrame(ctypes.Structure):
    _fields_ = [("rgb", ctypes.c_void_p), ("rgbStride", ctypes.c_uint32),
                ("pyramid", ctypes.c_void_p * 8), ("pyramidStride", ctypes.c_uint32 * 8),
                ("scaled", ctypes.c_void_p), ("scaledStride", ctypes.c_uint32),
                ("edges", ctypes.c_void_p), ("edgesStride", ctypes.c_uint32),
                ("corners", ctypes.c_void_p), ("numItems", ctypes.c_void_p), ("numCorners", ctypes.c_void_p)]


class FrameOutputs:
    """Input and output planes of one frame of the vision graph."""

    def __init__(self, w, h, dw, dh, levels, capacity):
        self.rgb = Plane(w, h, "RGB")
        self.scaled = Plane(dw, dh, "U8")
        self.pyr = [Plane(a, b, "U8") for (a, b) in pyramid_dims(w, h, levels)]
        self.edges = Plane(w, h, "U8")
        self.capacity = capacity
        self.corners = _dev_bytes(28 * max(capacity, 1))
        self.num_items = torch.zeros(1, dtype=torch.int32, device="cuda")
        self.count = torch.zeros(1, dtype=torch.int64, device="cuda")

    def results(self):
        n = int(self.num_items.item())
        kp = self.corners[:28 * n].cpu().numpy().view(KEYPOINT).copy()
        return dict(scaled=self.scaled.download(), pyramid=[p.download() for p in self.pyr],
                    edges=self.edges.download(), corners=kp, count=int(self.count.item()))


class FusedVisionGraph:
    """The vision graph over a batch of frames through mvx_VisionGraph_* (fused, frame-batched)."""

    def __init__(self, w, h, dw, dh, levels=5, lower=80, upper=100, k=0.04, thr=0.0005, dist=5.0, capacity=65536,
                 frames=1):
        K = kernels()
        K.mvx_VisionGraph_workspace_size.restype = ctypes.c_size_t
        self.params = VisionParams(w, h, dw, dh, levels, lower, upper, k, thr, dist, capacity)
        if not K.mvx_VisionGraph_supported(ctypes.byref(self.params)):
            raise ValueError("fused vision graph does not apply to these parameters")
        self.frames = [FrameOutputs(w, h, dw, dh, levels, capacity) for _ in range(frames)]
        arr = (VisionFrame * frames)()
        for i, f in enumerate(self.frames):
            d = arr[i]
            d.rgb, d.rgbStride = f.rgb.t.data_ptr(), f.rgb.stride
            for l, p in enumerate(f.pyr):
                d.pyramid[l], d.pyramidStride[l] = p.t.data_ptr(), p.stride
            d.scaled, d.scaledStride = f.scaled.t.data_ptr(), f.scaled.stride
            d.edges, d.edgesStride = f.edges.t.data_ptr(), f.edges.stride
            d.corners, d.numItems, d.numCorners = f.corners.data_ptr(), f.num_items.data_ptr(), f.count.data_ptr()
        self.ws_size = K.mvx_VisionGraph_workspace_size(ctypes.byref(self.params), frames)
        self.ws = _dev_bytes(self.ws_size)
        self._arr = arr
        rc = K.mvx_VisionGraph_prepare(_stream(), ctypes.byref(self.params), arr, frames, _ptr(self.ws),
                                       ctypes.c_size_t(self.ws_size))
        if rc != 0:
            raise RuntimeError(f"mvx_VisionGraph_prepare returned {rc}")

    def run(self):
        rc = kernels().mvx_VisionGraph_run(_stream(), ctypes.byref(self.params), len(self.frames), _ptr(self.ws),
                                           ctypes.c_size_t(self.ws_size))
        if rc != 0:
            raise RuntimeError(f"mvx_VisionGraph_run returned {rc}")


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class ImageLevel(ctypes.Structure):
    """mvx_image_level_t"""
    _fields_ = [("ptr", ctypes.c_void_p), ("width", ctypes.c_uint32), ("height", ctypes.c_uint32), ("stride", ctypes.c_uint32)]


class PyrLKTracker:
    """OpticalFlowPyrLK between two device pyramids (lists of Planes) with a persistent workspace
    (the reference node's local data); keypoints are device vx_keypoint_t arrays."""

    def __init__(self, old_levels, new_levels, capacity):
        K = kernels()
        K.mvx_OpticalFlowPyrLK_workspace_size.restype = ctypes.c_size_t
        self.old, self.new = old_levels, new_levels
        self.capacity = capacity
        self.ws_size = K.mvx_OpticalFlowPyrLK_workspace_size(old_levels[0].width, old_levels[0].height, capacity)
        self.ws = _dev_bytes(self.ws_size)
        self.out = _dev_bytes(28 * capacity)
        self.out_count = torch.zeros(1, dtype=torch.int32, device="cuda")
        n = len(old_levels)
        self.lo = (ImageLevel * n)(*[ImageLevel(p.t.data_ptr(), p.width, p.height, p.stride) for p in old_levels])
        self.ln = (ImageLevel * n)(*[ImageLevel(p.t.data_ptr(), p.width, p.height, p.stride) for p in new_levels])

    def set_points(self, old_kp, est_kp=None):
        self.n = len(old_kp)
        self.kp = torch.from_numpy(np.ascontiguousarray(old_kp, KEYPOINT).view(np.uint8).copy()).cuda()
        self.est = self.kp if est_kp is None else torch.from_numpy(np.ascontiguousarray(est_kp, KEYPOINT).view(np.uint8).copy()).cuda()
        self.count = torch.tensor([self.n], dtype=torch.int32, device="cuda")

    def run(self, termination=0xF002, epsilon=0.01, iterations=20, use_estimate=False, window=9, scale=0.5):
        call("OpticalFlowPyrLK_XY_XY", _ptr(self.out), _ptr(self.out_count), self.lo, self.ln, len(self.old), ctypes.c_float(scale),
             _ptr(self.kp), _ptr(self.count), _ptr(self.est), self.capacity, ctypes.c_int32(termination), ctypes.c_float(epsilon),
             iterations, ctypes.c_int32(int(use_estimate)), window, _ptr(self.ws), ctypes.c_size_t(self.ws_size))

    def result(self):
        return self.out[:28 * self.n].cpu().numpy().view(KEYPOINT).copy()


class ChainImage(ctypes.Structure):
    """mvx_chain_image_t"""
    _fields_ = [("src", ctypes.c_void_p), ("width", ctypes.c_uint32), ("height", ctypes.c_uint32),
                ("srcStride", ctypes.c_uint32),
                ("gx", ctypes.c_void_p), ("gxStride", ctypes.c_uint32),
                ("gy", ctypes.c_void_p), ("gyStride", ctypes.c_uint32),
                ("mag", ctypes.c_void_p), ("magStride", ctypes.c_uint32),
                ("phase", ctypes.c_void_p), ("phaseStride", ctypes.c_uint32),
                ("half", ctypes.c_void_p), ("halfStride", ctypes.c_uint32)]


class ChainPlan(ctypes.Structure):
    """mvx_chain_plan_t"""
    _fields_ = [("table", ctypes.c_void_p), ("count", ctypes.c_uint32), ("numTiles", ctypes.c_uint32),
                ("tile", ctypes.c_uint32), ("workgroups", ctypes.c_uint32),
                ("tileWidth", ctypes.c_uint32), ("tileHeight", ctypes.c_uint32)]


CHAIN_TILES = [(16, 16), (32, 16), (32, 32), (64, 16), (64, 32)]
CHAIN_OCCUPANCY_STEPS = [1, 2, 4, 0]


class ChainBatch:
    """Gaussian3x3 -> Sobel GXY -> Magnitude + Phase and ScaleImage to half size over a batch of U8
    device planes of uneven sizes: one persistent launch (mvx_ChainBatch_*), or the same work as
    plain per-node launches, image after image (``run_per_image``)."""

    def __init__(self, sources, with_half=True, virtual_gradients=False):
        """virtual_gradients: gx / gy are not stored and each image takes the merged SobelMagnitudePhase
        rules, as the reference's node-merge rules do when the gradients are virtual"""
        K = kernels()
        K.mvx_ChainBatch_workspace_size.restype = ctypes.c_size_t
        self.src = list(sources)
        self.virtual_gradients = virtual_gradients
        self.gx, self.gy, self.mag, self.phase, self.half, self.blur = [], [], [], [], [], []
        arr = (ChainImage * len(self.src))()
        for i, s in enumerate(self.src):
            w, h = s.width, s.height
            gx, gy = (None, None) if virtual_gradients else (Plane(w, h, "S16"), Plane(w, h, "S16"))
            mag, ph = Plane(w, h, "S16"), Plane(w, h, "U8")
            hf = Plane(w // 2, h // 2, "U8") if with_half else None
            self.gx.append(gx); self.gy.append(gy); self.mag.append(mag); self.phase.append(ph); self.half.append(hf)
            arr[i] = ChainImage(s.t.data_ptr(), w, h, s.stride, gx.t.data_ptr() if gx else None, gx.stride if gx else 0,
                                gy.t.data_ptr() if gy else None, gy.stride if gy else 0,
                                mag.t.data_ptr(), mag.stride, ph.t.data_ptr(), ph.stride,
                                hf.t.data_ptr() if hf else None, hf.stride if hf else 0)
        self._arr = arr
        self.ws_size = K.mvx_ChainBatch_workspace_size(len(self.src))
        self.ws = _dev_bytes(self.ws_size)
        self.plan = ChainPlan()
        self.graph_exec = ctypes.c_void_p(None)

    def prepare(self, tile=4, per_cu=0):
        rc = kernels().mvx_ChainBatch_prepare(self._arr, len(self.src), tile, per_cu, _ptr(self.ws),
                                              ctypes.c_size_t(self.ws_size), ctypes.byref(self.plan))
        if rc != 0:
            raise RuntimeError(f"mvx_ChainBatch_prepare returned {rc}")
        self.destroy_graph()
        return self

    def autotune(self, reps=3):
        """times every tile shape x workgroups-per-CU step; keeps the fastest plan; returns
        {(tile_w, tile_h, per_cu): ms}"""
        n = len(CHAIN_TILES) * len(CHAIN_OCCUPANCY_STEPS)
        table = (ctypes.c_float * n)()
        rc = kernels().mvx_ChainBatch_autotune(_stream(), self._arr, len(self.src), reps, _ptr(self.ws),
                                               ctypes.c_size_t(self.ws_size), ctypes.byref(self.plan), table)
        if rc != 0:
            raise RuntimeError(f"mvx_ChainBatch_autotune returned {rc}")
        self.destroy_graph()
        out = {}
        for t, (tw, th) in enumerate(CHAIN_TILES):
            for o, pc in enumerate(CHAIN_OCCUPANCY_STEPS):
                out[(tw, th, pc)] = table[t * len(CHAIN_OCCUPANCY_STEPS) + o]
        return out

    def run(self):
        rc = kernels().mvx_ChainBatch_run(_stream(), ctypes.byref(self.plan))
        if rc != 0:
            raise RuntimeError(f"mvx_ChainBatch_run returned {rc}")

    def capture(self):
        """instantiates a HIP graph holding the batched launch of the current plan"""
        self.destroy_graph()
        rc = kernels().mvx_ChainBatch_graph_create(ctypes.byref(self.plan), ctypes.byref(self.graph_exec))
        if rc != 0:
            raise RuntimeError(f"mvx_ChainBatch_graph_create returned {rc}")

    def launch_graph(self):
        rc = kernels().mvx_graph_launch(self.graph_exec, _stream())
        if rc != 0:
            raise RuntimeError(f"mvx_graph_launch returned {rc}")

    def destroy_graph(self):
        if self.graph_exec.value:
            kernels().mvx_graph_destroy(self.graph_exec)
            self.graph_exec = ctypes.c_void_p(None)

    def run_per_image(self):
        """the reference executor's model: one launch per node and image"""
        if self.virtual_gradients:
            raise NotImplementedError("per-image launches store the gradients")
        if not self.blur:
            self.blur = [Plane(s.width, s.height, "U8") for s in self.src]
        for s, b, gx, gy, mag, ph, hf in zip(self.src, self.blur, self.gx, self.gy, self.mag, self.phase, self.half):
            w, h = s.width, s.height
            call("Gaussian_U8_U8_3x3", w, h, b, b.stride, s, s.stride)
            call("Sobel_S16S16_U8_3x3_GXY", w, h, gx, gx.stride, gy, gy.stride, b, b.stride)
            call("Magnitude_S16_S16S16", w, h, mag, mag.stride, gx, gx.stride, gy, gy.stride)
            call("Phase_U8_S16S16", w, h, ph, ph.stride, gx, gx.stride, gy, gy.stride)
            if hf is not None:
                call("ScaleImage_U8_U8_Bilinear", w // 2, h // 2, hf, hf.stride, w, h, s, s.stride)

    def clear_outputs(self):
        for planes in (self.gx, self.gy, self.mag, self.phase, self.half):
            for p in planes:
                if p is not None:
                    p.t.fill_(0)

    def results(self, i):
        out = dict(mag=self.mag[i].download(), phase=self.phase[i].download())
        if not self.virtual_gradients:
            out.update(gx=self.gx[i].download(), gy=self.gy[i].download())
        if self.half[i] is not None:
            out["half"] = self.half[i].download()
        return out

    def __del__(self):
        try:
            self.destroy_graph()
        except Exception:
            pass


class LoomStitch:
    """The stitch variant of the uneven-batch configuration (synth.stitch_rig): cameras remapped into an
    equirectangular frame and merged with Loom's camera-selection maps, either in one pass
    (mvx_Stitch_Loom_RGB: each output pixel samples the cameras it selects) or in Loom's two passes
    (mvx_Remap_U24_U24_Bilinear_Constant per camera into a stacked frame, then mvx_StitchMerge_Loom_RGB)."""

    def __init__(self, rig):
        import torch
        self.W, self.H, self.sw, self.sh = rig["W"], rig["H"], rig["sw"], rig["sh"]
        self.count = len(rig["images"])
        W, H = self.W, self.H
        self.cams = [Plane.from_array(img, "RGB") for img in rig["images"]]
        self.fixed = torch.zeros((self.count * H, W * 4), dtype=torch.uint8, device="cuda")
        for c, m in enumerate(rig["maps"]):
            call("Remap_FixedMap_F32", W, H, ctypes.c_void_p(self.fixed[c * H].data_ptr()), W * 4, ctypes.c_void_p(m.data_ptr()), W * 8,
                 self.sw, self.sh)
        self.weights = rig["weights"].cuda()
        self.id0, self.id1, self.id2 = rig["camId0"].cuda(), rig["camId1"].cuda(), rig["camId2"].cuda()
        self.out = Plane(W, H, "RGB")
        self.stacked = None
        P, U = ctypes.c_void_p * self.count, ctypes.c_uint32 * self.count
        self._cam_ptrs = P(*[p.t.data_ptr() for p in self.cams])
        self._cam_strides = U(*[p.stride for p in self.cams])

    def _sel_args(self):
        g = self.W // 8
        return (ctypes.c_void_p(self.weights.data_ptr()), self.W, ctypes.c_void_p(self.id0.data_ptr()), g,
                ctypes.c_void_p(self.id1.data_ptr()), 2 * g, ctypes.c_void_p(self.id2.data_ptr()), 2 * g)

    def run_fused(self):
        call("Stitch_Loom_RGB", self.W, self.H, self.out, self.out.stride, self.count, self._cam_ptrs, self._cam_strides, self.sw,
             self.sh, ctypes.c_void_p(self.fixed.data_ptr()), self.W * 4, *self._sel_args())

    def run_two_pass(self):
        import torch
        W, H = self.W, self.H
        if self.stacked is None:
            self.stacked = torch.zeros((self.count * H, 3 * W), dtype=torch.uint8, device="cuda")
        for c, cam in enumerate(self.cams):
            call("Remap_U24_U24_Bilinear_Constant", W, H, ctypes.c_void_p(self.stacked[c * H].data_ptr()), 3 * W, self.sw, self.sh, cam,
                 cam.stride, ctypes.c_void_p(self.fixed[c * H].data_ptr()), W * 4, ctypes.c_uint8(0))
        call("StitchMerge_Loom_RGB", W, H, self.out, self.out.stride, self.count, ctypes.c_void_p(self.stacked.data_ptr()), 3 * W,
             *self._sel_args())

    def algorithmic_bytes(self):
        """bytes the one-pass stitch must move at least: the camera images once, every written output
        pixel (3 B), for every (pixel, sampled camera) its 4-byte table entry and, when blended, its
        weight byte, and the selection maps (1 B per 8-pixel group, plus 4 B for blended groups)"""
        import torch
        sel = self.id0.to(torch.int32)
        n = torch.where(sel == 31, 0, torch.where(sel < 31, 1, torch.where(sel > 131, 6, torch.where(sel > 128, sel - 126, 2))))
        # blended groups list up to n ids; ids of 31 are not sampled
        ids = torch.stack([(self.id1.to(torch.int32) >> (5 * k)) & 31 for k in range(3)] +
                          [(self.id2.to(torch.int32) >> (5 * k)) & 31 for k in range(3)])
        k = torch.arange(6, device=sel.device)[:, None, None]
        sampled_blend = ((ids < 31) & (k < n[None]) & (sel[None] > 31)).sum().item()
        single = (sel < 31).sum().item()
        written = (sel != 31).sum().item() * 8
        blend_groups = (sel > 31).sum().item()
        cams = sum(p.height * p.width * 3 for p in self.cams)
        return int(cams + 3 * written + 8 * (4 * single + 5 * sampled_blend) + sel.numel() + 4 * blend_groups)

    def results(self):
        return self.out.download()
