"""GPU parity of the Canny edge trace (mvx_CannyEdgeTrace_U8_U8XY) on its own: hysteresis maps of 0 /
127 / 255 in the interior -- random ones of several densities, long one-pixel paths that cross many
64x16 tiles, sizes on and off the tile grid, small images that take the per-pixel union-find -- against
the oracle's breadth-first flood, bit-exact."""
import ctypes

import numpy as np
import pytest
import torch

from mivisionx_amd.device import Plane, call
from util import assert_same

pytestmark = pytest.mark.gpu


def _trace(img):
    h, w = img.shape
    p = Plane.from_array(img)
    stack = torch.zeros(w * h * 4, dtype=torch.uint8, device="cuda")
    call("CannyEdgeTrace_U8_U8XY", w, h, p, p.stride, ctypes.c_void_p(stack.data_ptr()), w * h)
    return p.download()


def _random_map(w, h, p127, p255, seed):
    rng = np.random.default_rng(seed)
    u = rng.random((h, w))
    m = np.where(u < p255, 255, np.where(u < p255 + p127, 127, 0)).astype(np.uint8)
    m[0, :] = m[-1, :] = 0
    m[:, 0] = m[:, -1] = 0
    return m


@pytest.mark.parametrize("w,h", [(64, 32), (65, 33), (64, 16), (130, 17), (128, 96), (200, 150), (333, 77), (1920, 1080),
                                 (3840, 2160), (9, 7), (20, 5)])
@pytest.mark.parametrize("p127,p255", [(0.45, 0.002), (0.6, 0.0005), (0.3, 0.05)])
def test_trace_random_maps(gpu, orc, w, h, p127, p255):
    m = _random_map(w, h, p127, p255, w * 7 + h)
    print(f"trace {w}x{h} {p127}/{p255}", flush=True)
    assert_same(_trace(m), orc.canny_trace(m), f"trace {w}x{h} {p127}/{p255}")


@pytest.mark.parametrize("w,h", [(300, 200), (1000, 700)])
def test_trace_long_paths(gpu, orc, w, h):
    """a one-pixel serpentine of 127s through the whole image with a single 255 at its far end (every
    tile it crosses depends on the last one), and a second serpentine with no 255 (all of it -> 0)"""
    m = np.zeros((h, w), np.uint8)
    for y in range(2, h - 2, 4):
        m[y, 2:w - 2] = 127
        x = w - 3 if (y // 4) % 2 == 0 else 2
        if y + 4 < h - 2:
            m[y:y + 4, x] = 127
    m2 = m.copy()
    ys = np.nonzero(m[:, 2])[0]
    m[ys[-1], 2] = 255
    assert_same(_trace(m), orc.canny_trace(m), f"serpentine {w}x{h}")
    assert_same(_trace(m2), orc.canny_trace(m2), f"serpentine without a strong pixel {w}x{h}")


def test_trace_stack_alignment(gpu, orc):
    """a stack at any 4-byte aligned address takes the tiled trace (same result); one that is not
    4-byte aligned is refused before anything runs"""
    w, h = 640, 360
    m = _random_map(w, h, 0.45, 0.002, 5)
    want = orc.canny_trace(m)
    base = torch.zeros(w * h * 4 + 512, dtype=torch.uint8, device="cuda")
    for off in (0, 4, 64, 260):
        p = Plane.from_array(m)
        call("CannyEdgeTrace_U8_U8XY", w, h, p, p.stride, ctypes.c_void_p(base.data_ptr() + off), w * h)
        assert_same(p.download(), want, f"trace with the stack at offset {off}")
    from mivisionx_amd import kernels
    p = Plane.from_array(m)
    rc = kernels().mvx_CannyEdgeTrace_U8_U8XY(None, w, h, p.ptr, p.stride,
                                              ctypes.c_void_p(base.data_ptr() + 2), w * h)
    assert rc == -10
