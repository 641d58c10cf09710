"""GPU parity of every color_convert the reference runs (YUV -> RGB / RGBX, RGB / RGBX -> IYUV / NV12 /
YUV4, RGB <-> RGBX, the YUV -> YUV byte moves), through the OpenVX engine (vxColorConvertNode ->
the mvx_ColorConvert_* / mvx_FormatConvert_* kernels) and directly through the C ABI, against the
oracle, every plane bit-exact; and the luma-only form the engine runs for a virtual YUV image read
only by channel_extract(Y)."""
import ctypes

import numpy as np
import pytest
import torch

from pyoracle import COLOR_CONVERSIONS, color_planes
from util import assert_same

pytestmark = pytest.mark.gpu

SIZES = [(1920, 1080), (102, 6), (66, 4)]


def _planes(fmt, w, h, seed):
    rng = np.random.default_rng(seed)
    return [rng.integers(0, 256, shape, dtype=np.uint8) for shape in color_planes(fmt, w, h)]


@pytest.fixture(scope="module")
def vxc(gpu):
    from mivisionx_amd import openvx as vx
    L = vx.lib()
    ctx = L.vxCreateContext()
    yield vx, L, ctx
    cc = ctypes.c_void_p(ctx)
    L.vxReleaseContext(ctypes.byref(cc))


@pytest.mark.parametrize("src,dst", COLOR_CONVERSIONS)
def test_color_convert_node(vxc, orc, src, dst):
    vx, L, ctx = vxc
    for k, (w, h) in enumerate(SIZES):
        print(f"color_convert {src}->{dst} {w}x{h}", flush=True)
        planes = _planes(src, w, h, 31 * k + len(src) + 7 * len(dst))
        a, b = vx.create_image(ctx, w, h, src), vx.create_image(ctx, w, h, dst)
        vx.write_planes(a, src, w, h, planes)
        g = L.vxCreateGraph(ctx)
        assert L.vxColorConvertNode(g, a, b)
        vx.check(L.vxVerifyGraph(g), "vxVerifyGraph")
        vx.check(L.vxProcessGraph(g), "vxProcessGraph")
        got = vx.read_planes(b, dst, w, h)
        want = orc.color_convert(src, dst, w, h, planes)
        for p, (x, y) in enumerate(zip(got, want)):
            assert_same(x, y.reshape(x.shape), f"{src}->{dst} {w}x{h} plane {p}")
        gg = ctypes.c_void_p(g)
        L.vxReleaseGraph(ctypes.byref(gg))
        for im in (a, b):
            vx.release(im)


@pytest.mark.parametrize("src", ["RGB", "RGBX"])
@pytest.mark.parametrize("yuv", ["IYUV", "NV12", "YUV4"])
def test_luma_only_graph(vxc, orc, src, yuv):
    """RGB / RGBX -> virtual YUV -> channel_extract(Y): only the luma kernel runs, with its own rounding"""
    vx, L, ctx = vxc
    for w, h in ((1920, 1080), (102, 6), (37, 5)):
        planes = _planes(src, w, h, w + h)
        a, y = vx.create_image(ctx, w, h, src), vx.create_image(ctx, w, h, "U8")
        vx.write_planes(a, src, w, h, planes)
        g = L.vxCreateGraph(ctx)
        v = L.vxCreateVirtualImage(g, w, h, vx.DF[yuv])
        assert L.vxColorConvertNode(g, a, v) and L.vxChannelExtractNode(g, v, vx.VX_CHANNEL_Y, y)
        vx.check(L.vxVerifyGraph(g), "vxVerifyGraph")
        vx.check(L.vxProcessGraph(g), "vxProcessGraph")
        assert_same(vx.read_image(y), orc.color_luma(src, w, h, planes), f"{src}->{yuv} luma-only {w}x{h}")
        gg = ctypes.c_void_p(g)
        L.vxReleaseGraph(ctypes.byref(gg))
        for im in (a, y):
            vx.release(im)


def test_c_abi_direct(gpu, orc):
    """a few entry points called directly with padded (ALIGN16) planes, on a size with scalar tails"""
    from mivisionx_amd.device import Plane, call
    w, h = 1926, 54
    uyvy = _planes("UYVY", w, h, 1)[0]
    s = Plane(w, h, "U16")
    s.t[:, : 2 * w].copy_(torch.from_numpy(uyvy))
    d = Plane(w, h, "RGBX")
    call("ColorConvert_RGBX_UYVY", w, h, d, d.stride, s, s.stride)
    want = orc.color_convert("UYVY", "RGBX", w, h, [uyvy])[0]
    assert_same(d.t[:, : 4 * w].cpu().numpy(), want, "RGBX_UYVY")
    rgbx = _planes("RGBX", w, h, 2)[0]
    r = Plane(w, h, "RGBX")
    r.t[:, : 4 * w].copy_(torch.from_numpy(rgbx))
    Y, U, V = Plane(w, h), Plane(w, h), Plane(w, h)
    call("ColorConvert_YUV4_RGBX", w, h, Y, Y.stride, U, U.stride, V, V.stride, r, r.stride)
    for got, exp, k in zip((Y, U, V), orc.color_convert("RGBX", "YUV4", w, h, [rgbx]), "YUV"):
        assert_same(got.download(), exp, f"YUV4_RGBX {k}")
