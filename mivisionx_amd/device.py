"""Device image planes in the reference layout, and thin callers of the kernel C ABI.

A plane is a zero-initialised ``torch.uint8`` tensor of shape ``(height, stride)`` on the GPU with
``stride = ALIGN16(width * bytes_per_pixel)`` (ago_util.cpp:1107); padding bytes stay zero because
no kernel writes them.  ``call(name, ...)`` invokes an ``mvx_*`` entry point on the current torch
stream and raises on a non-zero status.
"""
import ctypes

import numpy as np
import torch

from . import kernels

BPP = {"U8": 1, "S16": 2, "U16": 2, "U32": 4, "S32": 4, "F32": 4, "RGB": 3, "RGBX": 4, "F32x3": 12}
NP_DTYPE = {"U8": np.uint8, "S16": np.int16, "U16": np.uint16, "U32": np.uint32, "S32": np.int32,
            "F32": np.float32, "RGB": np.uint8, "RGBX": np.uint8, "F32x3": np.float32}


def align16(x):
    return (x + 15) & ~15


class Plane:
    """One image plane in device memory with the reference's stride and zero padding."""

    def __init__(self, width, height, fmt="U8", device="cuda"):
        self.width, self.height, self.fmt = int(width), int(height), fmt
        self.bpp = BPP[fmt]
        self.stride = align16(self.width * self.bpp)
        self.t = torch.zeros((self.height, self.stride), dtype=torch.uint8, device=device)

    @property
    def ptr(self):
        return ctypes.c_void_p(self.t.data_ptr())

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        raw = arr.view(np.uint8).reshape(self.height, self.width * self.bpp)
        self.t[:, : self.width * self.bpp].copy_(torch.from_numpy(raw))
        return self

    def download(self):
        raw = self.t[:, : self.width * self.bpp].cpu().numpy()
        out = np.ascontiguousarray(raw).view(NP_DTYPE[self.fmt])
        if self.fmt in ("RGB", "RGBX"):
            return out.reshape(self.height, self.width, self.bpp)
        if self.fmt == "F32x3":
            return out.reshape(self.height, self.width, 3)
        return out.reshape(self.height, self.width)

    @classmethod
    def from_array(cls, arr, fmt="U8"):
        h, w = arr.shape[:2]
        return cls(w, h, fmt).upload(arr)


class MvxError(RuntimeError):
    pass


def stream_handle():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class KernelTimer:
    """Brackets every ``call`` with HIP events on the launch stream while active (``with timer:``);
    ``summary()`` gives per-entry-point launch counts and mean milliseconds."""

    active = None

    def __init__(self, only=None):
        self.only = only
        self.events = []

    def __enter__(self):
        KernelTimer.active = self
        return self

    def __exit__(self, *exc):
        KernelTimer.active = None

    def summary(self):
        torch.cuda.synchronize()
        out = {}
        for name, a, b in self.events:
            n, t = out.get(name, (0, 0.0))
            out[name] = (n + 1, t + a.elapsed_time(b))
        return {k: {"launches": n, "mean_ms": t / n, "total_ms": t} for k, (n, t) in out.items()}


def call(name, *args):
    """Call ``mvx_<name>(stream, *args)`` on the current torch stream; raise on failure."""
    fn = getattr(kernels(), "mvx_" + name)
    timer = KernelTimer.active
    if timer is not None and (timer.only is None or name in timer.only):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        rc = _invoke(fn, name, args)
        b.record()
        timer.events.append((name, a, b))
        return rc
    return _invoke(fn, name, args)


def _invoke(fn, name, args):
    conv = []
    for a in args:
        if isinstance(a, Plane):
            conv.append(a.ptr)
        elif isinstance(a, float):
            conv.append(ctypes.c_float(a))
        elif isinstance(a, int):
            conv.append(ctypes.c_uint32(a) if a >= 0 else ctypes.c_int32(a))
        else:
            conv.append(a)
    rc = fn(stream_handle(), *conv)
    if rc != 0:
        raise MvxError(f"mvx_{name} returned {rc}")
    return rc
