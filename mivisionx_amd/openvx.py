"""ctypes binding of libmvx_openvx.so: the vx*/vxu* C API of the engine, for tests, benchmarks and
Python callers.

Only declarations and data movement live here; every node runs inside the native engine on the GPU.
Constants are the OpenVX 1.3 values of include/VX/vx.h.
"""
import ctypes
import os

import numpy as np

from . import NativeLibraryMissing, kernels

_HERE = os.path.dirname(os.path.abspath(__file__))
OPENVX_SO = os.path.join(_HERE, "libmvx_openvx.so")
RUNVX = os.path.join(_HERE, "runvx")

c = ctypes
P = c.c_void_p
E = c.c_int32          # vx_enum / vx_status
U32 = c.c_uint32
SZ = c.c_size_t


def _df(a, b, cc, d):
    return ord(a) | (ord(b) << 8) | (ord(cc) << 16) | (ord(d) << 24)


def _enum(vendor_id, base):
    return (vendor_id << 20) | (base << 12)


VX_SUCCESS = 0
VX_ERROR_NOT_SUPPORTED = -3
VX_ERROR_INVALID_FORMAT = -14
VX_ERROR_INVALID_DIMENSION = -15
VX_ERROR_INVALID_VALUE = -16
VX_ERROR_INVALID_TYPE = -17
VX_ERROR_INVALID_GRAPH = -18
VX_ERROR_MULTIPLE_WRITERS = -23

DF = {"U8": _df("U", "0", "0", "8"), "S16": _df("S", "0", "1", "6"), "U16": _df("U", "0", "1", "6"),
      "RGB": _df("R", "G", "B", "2"), "RGBX": _df("R", "G", "B", "A"), "IYUV": _df("I", "Y", "U", "V"),
      "NV12": _df("N", "V", "1", "2"), "NV21": _df("N", "V", "2", "1"), "UYVY": _df("U", "Y", "V", "Y"),
      "YUYV": _df("Y", "U", "Y", "V"), "YUV4": _df("Y", "U", "V", "4"), "VIRT": _df("V", "I", "R", "T")}
BPP = {"U8": 1, "S16": 2, "U16": 2, "RGB": 3, "RGBX": 4}
NPT = {"U8": np.uint8, "S16": np.int16, "U16": np.uint16, "RGB": np.uint8, "RGBX": np.uint8}

VX_TYPE_INT32, VX_TYPE_FLOAT32, VX_TYPE_ENUM, VX_TYPE_SIZE = 0x006, 0x00A, 0x00C, 0x00D
VX_TYPE_UINT32 = 0x007
VX_TYPE_KEYPOINT = 0x021
VX_TYPE_COORDINATES2DF = 0x024
VX_READ_ONLY = _enum(0, 0x11) + 1
VX_WRITE_ONLY = _enum(0, 0x11) + 2
VX_MEMORY_TYPE_HOST = _enum(0, 0x0E) + 1
VX_MEMORY_TYPE_HIP = _enum(0, 0x0E) + 3
VX_INTERPOLATION_NEAREST_NEIGHBOR = _enum(0, 0x04) + 0
VX_INTERPOLATION_BILINEAR = _enum(0, 0x04) + 1
VX_INTERPOLATION_AREA = _enum(0, 0x04) + 2
VX_CHANNEL_R = _enum(0, 0x09) + 0x10
VX_CHANNEL_G = _enum(0, 0x09) + 0x11
VX_CHANNEL_B = _enum(0, 0x09) + 0x12
VX_CHANNEL_Y = _enum(0, 0x09) + 0x14
VX_CHANNEL_U = _enum(0, 0x09) + 0x15
VX_CHANNEL_V = _enum(0, 0x09) + 0x16
VX_THRESHOLD_TYPE_BINARY = _enum(0, 0x0B) + 0
VX_THRESHOLD_TYPE_RANGE = _enum(0, 0x0B) + 1
VX_NORM_L1 = _enum(0, 0x10) + 0
VX_NORM_L2 = _enum(0, 0x10) + 1
VX_BORDER_UNDEFINED = _enum(0, 0x0C) + 0
VX_BORDER_CONSTANT = _enum(0, 0x0C) + 1
VX_BORDER_REPLICATE = _enum(0, 0x0C) + 2
VX_TERM_CRITERIA_ITERATIONS = _enum(0, 0x0F) + 0
VX_TERM_CRITERIA_EPSILON = _enum(0, 0x0F) + 1
VX_TERM_CRITERIA_BOTH = _enum(0, 0x0F) + 2
VX_SCALE_PYRAMID_HALF = 0.5


def _attr(vendor, obj):
    return (vendor << 20) | (obj << 8)


VX_TYPE_GRAPH, VX_TYPE_NODE, VX_TYPE_IMAGE, VX_TYPE_ARRAY, VX_TYPE_CONTEXT = 0x802, 0x803, 0x80F, 0x80E, 0x801
VX_TYPE_SCALAR, VX_TYPE_THRESHOLD = 0x80D, 0x80A
VX_GRAPH_NUMNODES = _attr(0, VX_TYPE_GRAPH) + 0
VX_GRAPH_PERFORMANCE = _attr(0, VX_TYPE_GRAPH) + 2
VX_GRAPH_STATE = _attr(0, VX_TYPE_GRAPH) + 4
VX_NODE_STATUS = _attr(0, VX_TYPE_NODE) + 0
VX_NODE_BORDER = _attr(0, VX_TYPE_NODE) + 2
VX_IMAGE_WIDTH = _attr(0, VX_TYPE_IMAGE) + 0
VX_IMAGE_HEIGHT = _attr(0, VX_TYPE_IMAGE) + 1
VX_IMAGE_FORMAT = _attr(0, VX_TYPE_IMAGE) + 2
VX_ARRAY_NUMITEMS = _attr(0, VX_TYPE_ARRAY) + 1
VX_THRESHOLD_THRESHOLD_LOWER = _attr(0, VX_TYPE_THRESHOLD) + 2
VX_THRESHOLD_THRESHOLD_UPPER = _attr(0, VX_TYPE_THRESHOLD) + 3
VX_CONTEXT_ATTRIBUTE_AMD_AFFINITY = _attr(0x00D, VX_TYPE_CONTEXT) + 2
VX_GRAPH_ATTRIBUTE_AMD_AFFINITY = _attr(0x00D, VX_TYPE_GRAPH) + 1
VX_GRAPH_ATTRIBUTE_AMD_IMPORT_FROM_TEXT = _attr(0x00D, VX_TYPE_GRAPH) + 2
VX_GRAPH_ATTRIBUTE_AMD_OPTIMIZER_FLAGS = _attr(0x00D, VX_TYPE_GRAPH) + 4
VX_NODE_ATTRIBUTE_AMD_AFFINITY = _attr(0x00D, VX_TYPE_NODE) + 1
VX_IMAGE_ATTRIBUTE_AMD_HIP_BUFFER = _attr(0x00D, VX_TYPE_IMAGE) + 6
AGO_TARGET_AFFINITY_CPU, AGO_TARGET_AFFINITY_GPU = 0x10, 0x20
AGO_GRAPH_OPTIMIZER_FLAG_NO_SUPERNODE_MERGE = 0x20
AGO_GRAPH_OPTIMIZER_FLAG_NO_NODE_MERGE = 0x08

KEYPOINT = np.dtype([("x", np.int32), ("y", np.int32), ("strength", np.float32), ("scale", np.float32),
                     ("orientation", np.float32), ("tracking_status", np.int32), ("error", np.float32)])


class Rect(c.Structure):
    _fields_ = [("start_x", U32), ("start_y", U32), ("end_x", U32), ("end_y", U32)]


class Addressing(c.Structure):
    _fields_ = [("dim_x", U32), ("dim_y", U32), ("stride_x", c.c_int32), ("stride_y", c.c_int32),
                ("scale_x", U32), ("scale_y", U32), ("step_x", U32), ("step_y", c.c_uint16),
                ("stride_x_bits", c.c_uint16)]


class PixelValue(c.Union):
    _fields_ = [("U8", c.c_uint8), ("U16", c.c_uint16), ("S16", c.c_int16), ("U32", c.c_uint32),
                ("S32", c.c_int32), ("RGB", c.c_uint8 * 3), ("reserved", c.c_uint8 * 16)]


class Border(c.Structure):
    _fields_ = [("mode", E), ("constant_value", PixelValue)]


class Perf(c.Structure):
    _fields_ = [(n, c.c_uint64) for n in ("tmp", "beg", "end", "sum", "avg", "min", "num", "max")]


class Affinity(c.Structure):
    _fields_ = [("device_type", U32), ("device_info", U32), ("group", U32), ("reserved", U32)]


REGISTRY_CB = c.CFUNCTYPE(None, P, P, c.c_char_p, c.c_char_p)
LOG_CB = c.CFUNCTYPE(None, P, P, E, c.c_char_p)


class ImportInfo(c.Structure):
    _fields_ = [("text", c.c_char_p), ("num_ref", U32), ("ref", c.POINTER(P)), ("dumpToConsole", c.c_int32),
                ("data_registry_callback_f", REGISTRY_CB), ("data_registry_callback_obj", P)]


_PROTOS = {
    "vxCreateContext": (P, []), "vxReleaseContext": (E, [c.POINTER(P)]), "vxGetStatus": (E, [P]),
    "vxQueryContext": (E, [P, E, P, SZ]), "vxSetContextAttribute": (E, [P, E, P, SZ]),
    "vxQueryReference": (E, [P, E, P, SZ]), "vxReleaseReference": (E, [c.POINTER(P)]),
    "vxSetReferenceName": (E, [P, c.c_char_p]),
    "vxGetKernelByName": (P, [P, c.c_char_p]), "vxGetKernelByEnum": (P, [P, E]),
    "vxCreateImage": (P, [P, U32, U32, U32]), "vxCreateVirtualImage": (P, [P, U32, U32, U32]),
    "vxCreateUniformImage": (P, [P, U32, U32, U32, c.POINTER(PixelValue)]),
    "vxCreateImageFromHandle": (P, [P, U32, c.POINTER(Addressing), c.POINTER(P), E]),
    "vxQueryImage": (E, [P, E, P, SZ]), "vxSetImageAttribute": (E, [P, E, P, SZ]),
    "vxReleaseImage": (E, [c.POINTER(P)]),
    "vxCopyImagePatch": (E, [P, c.POINTER(Rect), U32, c.POINTER(Addressing), P, E, E]),
    "vxMapImagePatch": (E, [P, c.POINTER(Rect), U32, c.POINTER(c.c_size_t), c.POINTER(Addressing), c.POINTER(P), E, E, U32]),
    "vxUnmapImagePatch": (E, [P, c.c_size_t]),
    "vxCreateScalar": (P, [P, E, P]), "vxCopyScalar": (E, [P, P, E, E]), "vxReleaseScalar": (E, [c.POINTER(P)]),
    "vxCreateThreshold": (P, [P, E, E]), "vxCreateThresholdForImage": (P, [P, E, U32, U32]),
    "vxSetThresholdAttribute": (E, [P, E, P, SZ]), "vxCopyThresholdRange": (E, [P, c.POINTER(PixelValue), c.POINTER(PixelValue), E, E]),
    "vxReleaseThreshold": (E, [c.POINTER(P)]),
    "vxCreateArray": (P, [P, E, SZ]), "vxQueryArray": (E, [P, E, P, SZ]), "vxAddArrayItems": (E, [P, SZ, P, SZ]),
    "vxCopyArrayRange": (E, [P, SZ, SZ, SZ, P, E, E]), "vxTruncateArray": (E, [P, SZ]), "vxReleaseArray": (E, [c.POINTER(P)]),
    "vxCreatePyramid": (P, [P, SZ, c.c_float, U32, U32, U32]), "vxCreateVirtualPyramid": (P, [P, SZ, c.c_float, U32, U32, U32]),
    "vxGetPyramidLevel": (P, [P, U32]), "vxReleasePyramid": (E, [c.POINTER(P)]),
    "vxCreateMatrix": (P, [P, E, SZ, SZ]), "vxCopyMatrix": (E, [P, P, E, E]), "vxReleaseMatrix": (E, [c.POINTER(P)]),
    "vxCreateRemap": (P, [P, U32, U32, U32, U32]), "vxCopyRemapPatch": (E, [P, c.POINTER(Rect), SZ, P, E, E, E]),
    "vxReleaseRemap": (E, [c.POINTER(P)]),
    "vxCreateGraph": (P, [P]), "vxReleaseGraph": (E, [c.POINTER(P)]), "vxVerifyGraph": (E, [P]),
    "vxProcessGraph": (E, [P]), "vxScheduleGraph": (E, [P]), "vxWaitGraph": (E, [P]),
    "vxQueryGraph": (E, [P, E, P, SZ]), "vxSetGraphAttribute": (E, [P, E, P, SZ]), "vxIsGraphVerified": (E, [P]),
    "vxQueryNode": (E, [P, E, P, SZ]), "vxSetNodeAttribute": (E, [P, E, P, SZ]), "vxReleaseNode": (E, [c.POINTER(P)]),
    "vxRegisterLogCallback": (None, [P, LOG_CB, E]),
    "vxCreateGenericNode": (P, [P, P]), "vxSetParameterByIndex": (E, [P, U32, P]),
    "vxColorConvertNode": (P, [P, P, P]), "vxChannelExtractNode": (P, [P, P, E, P]),
    "vxuColorConvert": (E, [P, P, P]), "vxuChannelExtract": (E, [P, P, E, P]),
    "vxSobel3x3Node": (P, [P, P, P, P]), "vxMagnitudeNode": (P, [P, P, P, P]), "vxPhaseNode": (P, [P, P, P, P]),
    "vxScaleImageNode": (P, [P, P, P, E]), "vxBox3x3Node": (P, [P, P, P]), "vxGaussian3x3Node": (P, [P, P, P]),
    "vxMedian3x3Node": (P, [P, P, P]), "vxGaussianPyramidNode": (P, [P, P, P]),
    "vxCannyEdgeDetectorNode": (P, [P, P, P, c.c_int32, E, P]),
    "vxHarrisCornersNode": (P, [P, P, P, P, P, c.c_int32, c.c_int32, P, P]),
    "vxFastCornersNode": (P, [P, P, P, E, P, P]), "vxuFastCorners": (E, [P, P, P, E, P, P]),
    "vxWarpAffineNode": (P, [P, P, P, E, P]), "vxWarpPerspectiveNode": (P, [P, P, P, E, P]),
    "vxRemapNode": (P, [P, P, P, E, P]), "vxHalfScaleGaussianNode": (P, [P, P, P, c.c_int32]),
    "vxOpticalFlowPyrLKNode": (P, [P, P, P, P, P, P, E, P, P, P, SZ]),
    "vxuGaussian3x3": (E, [P, P, P]), "vxuBox3x3": (E, [P, P, P]), "vxuMedian3x3": (E, [P, P, P]),
    "vxuSobel3x3": (E, [P, P, P, P]), "vxuMagnitude": (E, [P, P, P, P]), "vxuPhase": (E, [P, P, P, P]),
    "vxuScaleImage": (E, [P, P, P, E]), "vxuGaussianPyramid": (E, [P, P, P]),
    "vxuCannyEdgeDetector": (E, [P, P, P, c.c_int32, E, P]),
    "vxuHarrisCorners": (E, [P, P, P, P, P, c.c_int32, c.c_int32, P, P]),
    "vxuWarpAffine": (E, [P, P, P, E, P]), "vxuWarpPerspective": (E, [P, P, P, E, P]), "vxuRemap": (E, [P, P, P, E, P]),
    "vxuHalfScaleGaussian": (E, [P, P, P, c.c_int32]),
    "vxuOpticalFlowPyrLK": (E, [P, P, P, P, P, P, E, P, P, P, SZ]),
}

_lib = None


def lib():
    """The loaded libmvx_openvx.so with prototypes declared (raises NativeLibraryMissing)."""
    global _lib
    if _lib is None:
        if not os.path.exists(OPENVX_SO):
            raise NativeLibraryMissing(f"{OPENVX_SO} is not built; run `python -c 'import __graft_entry__ as g; g.build()'`")
        kernels()
        L = c.CDLL(OPENVX_SO)
        for name, (res, args) in _PROTOS.items():
            fn = getattr(L, name)
            fn.restype, fn.argtypes = res, args
        _lib = L
    return _lib


class VXError(RuntimeError):
    def __init__(self, what, status):
        super().__init__(f"{what} returned {status}")
        self.status = status


def check(status, what="vx call"):
    if status != VX_SUCCESS:
        raise VXError(what, status)
    return status


def _ok(ref, what):
    if not ref:
        raise VXError(what, "NULL")
    return ref


def release(ref):
    h = P(ref)
    return lib().vxReleaseReference(c.byref(h))


def create_image(ctx, w, h, fmt):
    return _ok(lib().vxCreateImage(ctx, w, h, DF[fmt]), "vxCreateImage")


def write_image(img, arr, plane=0, sub=1):
    """copy a host array (h, w[, channels]) into plane `plane` of an image; `sub` is the plane's
    subsampling (2 for the chroma planes of IYUV / NV12: the rectangle is in image coordinates)"""
    arr = np.ascontiguousarray(arr)
    h, w = arr.shape[:2]
    bpp = arr.dtype.itemsize * (arr.shape[2] if arr.ndim == 3 else 1)
    addr = Addressing(w, h, bpp, w * bpp, 1024, 1024, 1, 1, 0)
    check(lib().vxCopyImagePatch(img, c.byref(Rect(0, 0, w * sub, h * sub)), plane, c.byref(addr), arr.ctypes.data_as(P),
                                 VX_WRITE_ONLY, VX_MEMORY_TYPE_HOST), "vxCopyImagePatch(write)")


def _plane_layout(fmt, w, h):
    """(samples across, rows, bytes per sample, subsampling) of each plane of a w x h image of `fmt`"""
    return {"RGB": [(w, h, 3, 1)], "RGBX": [(w, h, 4, 1)], "UYVY": [(w, h, 2, 1)], "YUYV": [(w, h, 2, 1)],
            "IYUV": [(w, h, 1, 1), (w // 2, h // 2, 1, 2), (w // 2, h // 2, 1, 2)],
            "NV12": [(w, h, 1, 1), (w // 2, h // 2, 2, 2)], "NV21": [(w, h, 1, 1), (w // 2, h // 2, 2, 2)],
            "YUV4": [(w, h, 1, 1)] * 3, "U8": [(w, h, 1, 1)]}[fmt]


def write_planes(img, fmt, w, h, planes):
    """every plane of an image from tightly packed host arrays (rows of samples x bytes per sample)"""
    for k, ((pw, ph, bpp, sub), arr) in enumerate(zip(_plane_layout(fmt, w, h), planes)):
        arr = np.ascontiguousarray(arr, dtype=np.uint8).reshape(ph, pw * bpp)
        addr = Addressing(w, h, bpp, pw * bpp, 1024, 1024, sub, sub, 0)
        check(lib().vxCopyImagePatch(img, c.byref(Rect(0, 0, w, h)), k, c.byref(addr), arr.ctypes.data_as(P), VX_WRITE_ONLY,
                                     VX_MEMORY_TYPE_HOST), "vxCopyImagePatch(write)")


def read_planes(img, fmt, w, h):
    """every plane of an image as tightly packed host arrays (rows of samples x bytes per sample)"""
    out = []
    for k, (pw, ph, bpp, sub) in enumerate(_plane_layout(fmt, w, h)):
        arr = np.zeros((ph, pw * bpp), np.uint8)
        addr = Addressing(w, h, bpp, pw * bpp, 1024, 1024, sub, sub, 0)
        check(lib().vxCopyImagePatch(img, c.byref(Rect(0, 0, w, h)), k, c.byref(addr), arr.ctypes.data_as(P), VX_READ_ONLY,
                                     VX_MEMORY_TYPE_HOST), "vxCopyImagePatch(read)")
        out.append(arr)
    return out


def read_image(img, fmt=None, plane=0):
    L = lib()
    w, h, f = U32(), U32(), U32()
    L.vxQueryImage(img, VX_IMAGE_WIDTH, c.byref(w), 4)
    L.vxQueryImage(img, VX_IMAGE_HEIGHT, c.byref(h), 4)
    L.vxQueryImage(img, VX_IMAGE_FORMAT, c.byref(f), 4)
    if fmt is None:
        fmt = {v: k for k, v in DF.items()}[f.value]
    bpp = BPP[fmt]
    out = np.zeros((h.value, w.value * bpp), np.uint8)
    addr = Addressing(w.value, h.value, bpp, w.value * bpp, 1024, 1024, 1, 1, 0)
    check(L.vxCopyImagePatch(img, c.byref(Rect(0, 0, w.value, h.value)), plane, c.byref(addr), out.ctypes.data_as(P),
                             VX_READ_ONLY, VX_MEMORY_TYPE_HOST), "vxCopyImagePatch(read)")
    out = out.view(NPT[fmt])
    return out.reshape(h.value, w.value, bpp) if fmt in ("RGB", "RGBX") else out.reshape(h.value, w.value)


def scalar(ctx, vx_type, value):
    ct = {VX_TYPE_INT32: c.c_int32, VX_TYPE_UINT32: c.c_uint32, VX_TYPE_FLOAT32: c.c_float, VX_TYPE_ENUM: c.c_int32,
          VX_TYPE_SIZE: c.c_size_t}[vx_type]
    v = ct(value)
    return _ok(lib().vxCreateScalar(ctx, vx_type, c.byref(v)), "vxCreateScalar")


def read_scalar(s, vx_type):
    ct = {VX_TYPE_INT32: c.c_int32, VX_TYPE_UINT32: c.c_uint32, VX_TYPE_FLOAT32: c.c_float, VX_TYPE_ENUM: c.c_int32,
          VX_TYPE_SIZE: c.c_size_t}[vx_type]
    v = ct()
    check(lib().vxCopyScalar(s, c.byref(v), VX_READ_ONLY, VX_MEMORY_TYPE_HOST), "vxCopyScalar")
    return v.value


def range_threshold(ctx, lower, upper, fmt="U8"):
    """a RANGE threshold for `fmt` input images ("U8" or "S16", whose values may exceed 255)"""
    t = _ok(lib().vxCreateThresholdForImage(ctx, VX_THRESHOLD_TYPE_RANGE, DF[fmt], DF["U8"]), "vxCreateThresholdForImage")
    lo, hi = PixelValue(), PixelValue()
    setattr(lo, fmt, lower)
    setattr(hi, fmt, upper)
    check(lib().vxCopyThresholdRange(t, c.byref(lo), c.byref(hi), VX_WRITE_ONLY, VX_MEMORY_TYPE_HOST), "vxCopyThresholdRange")
    return t


def threshold_range(t, fmt="U8"):
    lo, hi = PixelValue(), PixelValue()
    check(lib().vxCopyThresholdRange(t, c.byref(lo), c.byref(hi), VX_READ_ONLY, VX_MEMORY_TYPE_HOST), "vxCopyThresholdRange")
    return getattr(lo, fmt), getattr(hi, fmt)


def read_array(arr, dtype):
    n = c.c_size_t()
    check(lib().vxQueryArray(arr, VX_ARRAY_NUMITEMS, c.byref(n), c.sizeof(n)), "vxQueryArray")
    out = np.zeros(n.value, dtype)
    if n.value:
        check(lib().vxCopyArrayRange(arr, 0, n.value, dtype.itemsize, out.ctypes.data_as(P), VX_READ_ONLY,
                                     VX_MEMORY_TYPE_HOST), "vxCopyArrayRange")
    return out


def import_gdf(graph, text, refs=(), callback=None):
    """vxSetGraphAttribute(IMPORT_FROM_TEXT); `callback(ref, name, params)` receives runvx I/O fields"""
    arr = (P * max(1, len(refs)))(*refs)
    cb = REGISTRY_CB((lambda obj, ref, name, params: callback(ref, name.decode(), params.decode())) if callback
                     else (lambda *a: None))
    info = ImportInfo(text.encode(), len(refs), arr, 0, cb, None)
    status = lib().vxSetGraphAttribute(graph, VX_GRAPH_ATTRIBUTE_AMD_IMPORT_FROM_TEXT, c.byref(info), c.sizeof(info))
    return status
