"""GPU parity of YUV -> RGB colour conversion (IYUV and NV12) through the C ABI and the color_convert
node, and of channel_extract U / V from IYUV, against the oracle, bit-exact."""
import ctypes

import numpy as np
import pytest
import torch

from mivisionx_amd.device import Plane, call
from util import assert_same

pytestmark = pytest.mark.gpu


def _planes(w, h, seed):
    rng = np.random.default_rng(seed)
    return (rng.integers(0, 256, (h, w), dtype=np.uint8), rng.integers(0, 256, (h // 2, w // 2), dtype=np.uint8),
            rng.integers(0, 256, (h // 2, w // 2), dtype=np.uint8))


@pytest.mark.parametrize("w,h", [(64, 48), (100, 60), (38, 30), (1920, 1080), (16, 2)])
def test_rgb_from_yuv_c_abi(gpu, orc, w, h):
    y, u, v = _planes(w, h, w * h)
    Y, U, V = Plane.from_array(y), Plane.from_array(u), Plane.from_array(v)
    d = Plane(w, h, "RGB")
    call("ColorConvert_RGB_IYUV", w, h, d, d.stride, Y, Y.stride, U, U.stride, V, V.stride)
    assert_same(d.download(), orc.color_rgb_yuv(y, u, v), f"IYUV->RGB {w}x{h}")
    uv = np.stack([u, v], -1).reshape(h // 2, w)
    UV = Plane.from_array(uv)
    d = Plane(w, h, "RGB")
    call("ColorConvert_RGB_NV12", w, h, d, d.stride, Y, Y.stride, UV, UV.stride)
    assert_same(d.download(), orc.color_rgb_yuv(y, uv), f"NV12->RGB {w}x{h}")


def test_yuv_nodes(gpu, orc):
    from mivisionx_amd import openvx as vx
    L = vx.lib()
    ctx = L.vxCreateContext()
    try:
        w, h = 1280, 720
        y, u, v = _planes(w, h, 9)
        iyuv = vx.create_image(ctx, w, h, "IYUV")
        for k, p in enumerate((y, u, v)):
            vx.write_image(iyuv, p, plane=k, sub=1 if k == 0 else 2)
        rgb = vx.create_image(ctx, w, h, "RGB")
        assert L.vxuColorConvert(ctx, iyuv, rgb) == 0
        assert_same(vx.read_image(rgb), orc.color_rgb_yuv(y, u, v), "vxuColorConvert IYUV->RGB")
        for ch, want in ((vx.VX_CHANNEL_U, u), (vx.VX_CHANNEL_V, v), (vx.VX_CHANNEL_Y, y)):
            out = vx.create_image(ctx, want.shape[1], want.shape[0], "U8")
            assert L.vxuChannelExtract(ctx, iyuv, ch, out) == 0
            assert_same(vx.read_image(out), want, f"channel_extract {ch:#x}")
        # RGB channels
        for ch, k in ((vx.VX_CHANNEL_R, 0), (vx.VX_CHANNEL_G, 1), (vx.VX_CHANNEL_B, 2)):
            out = vx.create_image(ctx, w, h, "U8")
            assert L.vxuChannelExtract(ctx, rgb, ch, out) == 0
            assert_same(vx.read_image(out), vx.read_image(rgb)[..., k], f"RGB channel_extract {k}")
        # NV12: the interleaved chroma plane, U at byte 0 and V at byte 1 of each pair
        nv12 = vx.create_image(ctx, w, h, "NV12")
        vx.write_image(nv12, y, plane=0)
        vx.write_image(nv12, np.stack([u, v], -1).reshape(h // 2, w // 2, 2), plane=1, sub=2)
        assert L.vxuColorConvert(ctx, nv12, rgb) == 0
        assert_same(vx.read_image(rgb), orc.color_rgb_yuv(y, np.stack([u, v], -1).reshape(h // 2, w)), "vxuColorConvert NV12->RGB")
        for ch, want in ((vx.VX_CHANNEL_U, u), (vx.VX_CHANNEL_V, v)):
            out = vx.create_image(ctx, want.shape[1], want.shape[0], "U8")
            assert L.vxuChannelExtract(ctx, nv12, ch, out) == 0
            assert_same(vx.read_image(out), want, f"NV12 channel_extract {ch:#x}")
    finally:
        cc = ctypes.c_void_p(ctx)
        L.vxReleaseContext(ctypes.byref(cc))
