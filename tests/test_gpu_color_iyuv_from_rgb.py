"""GPU parity of RGB -> IYUV colour conversion, all three planes, through the C ABI and the color_convert
node (non-virtual IYUV output), against the oracle, bit-exact."""
import ctypes

import numpy as np
import pytest

from mivisionx_amd.device import Plane, call
from util import assert_same

pytestmark = pytest.mark.gpu


def _rgb(w, h, seed):
    return np.random.default_rng(seed).integers(0, 256, (h, w, 3), dtype=np.uint8)


@pytest.mark.parametrize("w,h", [(64, 48), (38, 30), (6, 4), (2, 2), (130, 10), (1920, 1080), (2050, 6)])
def test_iyuv_from_rgb_c_abi(gpu, orc, w, h):
    rgb = _rgb(w, h, w + 3 * h)
    S = Plane.from_array(rgb, "RGB")
    Y, U, V = Plane(w, h, "U8"), Plane(w // 2, h // 2, "U8"), Plane(w // 2, h // 2, "U8")
    call("ColorConvert_IYUV_RGB", w, h, Y, Y.stride, U, U.stride, V, V.stride, S, S.stride)
    for got, want, name in zip((Y, U, V), orc.color_iyuv_from_rgb(rgb), "YUV"):
        assert_same(got.download(), want, f"RGB->IYUV {name} {w}x{h}")


def test_iyuv_from_rgb_node(gpu, orc):
    from mivisionx_amd import openvx as vx
    L = vx.lib()
    ctx = L.vxCreateContext()
    try:
        w, h = 1278, 720
        rgb = _rgb(w, h, 5)
        src = vx.create_image(ctx, w, h, "RGB")
        vx.write_image(src, rgb)
        iyuv = vx.create_image(ctx, w, h, "IYUV")
        assert L.vxuColorConvert(ctx, src, iyuv) == 0
        for ch, want in zip((vx.VX_CHANNEL_Y, vx.VX_CHANNEL_U, vx.VX_CHANNEL_V), orc.color_iyuv_from_rgb(rgb)):
            out = vx.create_image(ctx, want.shape[1], want.shape[0], "U8")
            assert L.vxuChannelExtract(ctx, iyuv, ch, out) == 0
            assert_same(vx.read_image(out), want, f"vxuColorConvert RGB->IYUV channel {ch:#x}")
        # RGB -> NV12 through vxu: the luma plane and the interleaved chroma of ColorConvert_NV12_RGB
        nv12 = vx.create_image(ctx, w, h, "NV12")
        assert L.vxuColorConvert(ctx, src, nv12) == 0
        want = orc.color_convert("RGB", "NV12", w, h, [rgb.reshape(h, 3 * w)])
        for k, (got, exp) in enumerate(zip(vx.read_planes(nv12, "NV12", w, h), want)):
            assert_same(got, exp, f"vxuColorConvert RGB->NV12 plane {k}")
    finally:
        cc = ctypes.c_void_p(ctx)
        L.vxReleaseContext(ctypes.byref(cc))


@pytest.mark.parametrize("readers", ["YUV", "UV", "Y"])
def test_virtual_iyuv_chroma_readers_get_chroma(gpu, orc, readers):
    """RGB -> virtual IYUV -> channel_extract: the chroma kernels run whenever a U or V channel is read
    (the reference divides RGB -> IYUV into Y, IU and IV kernels and keeps the ones whose outputs are read)"""
    from mivisionx_amd import openvx as vx
    L = vx.lib()
    ctx = L.vxCreateContext()
    try:
        w, h = 334, 120
        rgb = _rgb(w, h, 9)
        g = L.vxCreateGraph(ctx)
        src = vx.create_image(ctx, w, h, "RGB")
        vx.write_image(src, rgb)
        yuv = L.vxCreateVirtualImage(g, w, h, vx.DF["IYUV"])
        assert L.vxColorConvertNode(g, src, yuv)
        want = dict(zip("YUV", orc.color_iyuv_from_rgb(rgb)))
        chan = dict(Y=vx.VX_CHANNEL_Y, U=vx.VX_CHANNEL_U, V=vx.VX_CHANNEL_V)
        outs = {}
        for k in readers:
            outs[k] = vx.create_image(ctx, want[k].shape[1], want[k].shape[0], "U8")
            assert L.vxChannelExtractNode(g, yuv, chan[k], outs[k])
        assert L.vxProcessGraph(g) == 0
        for k, o in outs.items():
            assert_same(vx.read_image(o), want[k], f"virtual IYUV channel {k} (readers {readers})")
    finally:
        cc = ctypes.c_void_p(ctx)
        L.vxReleaseContext(ctypes.byref(cc))
