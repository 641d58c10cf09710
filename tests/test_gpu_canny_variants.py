"""GPU parity of Canny with gradient sizes 5 and 7 and the L2 norm: the U16 gradient planes through the
C ABI (padding included, so the slot the narrow-image loop writes before each row is compared) and
the whole canny_edge_detector node through the OpenVX API, including the gradient-7 stage's in-place
division of the threshold by 4 on every run -- against the oracle, bit-exact."""
import ctypes

import numpy as np
import pytest
import torch

from mivisionx_amd import synth
from mivisionx_amd.device import Plane, call
from util import assert_same

pytestmark = pytest.mark.gpu

VARIANTS = [(5, 0), (7, 0), (3, 1), (5, 1), (7, 1)]
THR = {3: (40, 110), 5: (400, 1300), 7: (1600, 5000)}


def _src(w, h, seed=7):
    return synth.scene_u8(w, h, seed) if w >= 16 and h >= 16 else synth.uniform_u8(w, h, seed)


@pytest.mark.parametrize("w,h", [(64, 48), (100, 60), (37, 29), (12, 10), (15, 9), (333, 77), (1920, 1080)])
@pytest.mark.parametrize("g,l2", VARIANTS)
def test_canny_sobel_planes(gpu, orc, w, h, g, l2):
    src = _src(w, h)
    s = Plane.from_array(src)
    d = Plane(w, h, "U16")
    call(f"CannySobel_U16_U8_{g}x{g}_{'L2' if l2 else 'L1'}NORM", w, h, d, d.stride, s, s.stride)
    got = d.t.cpu().numpy().view(np.uint16)
    assert_same(got, orc.canny_sobel_ex(src, g, l2, padded=True), f"canny sobel {g} l2={l2} {w}x{h}")


@pytest.mark.parametrize("w,h", [(3840, 24), (16, 20), (977, 33), (1024, 17), (124, 9)])
@pytest.mark.parametrize("g,l2", VARIANTS)
def test_canny_sobel_planes_noise(gpu, orc, w, h, g, l2):
    """uniform and 0/255 noise (the largest gradients), widths at and between the kernel's column strips"""
    for k, src in enumerate((synth.uniform_u8(w, h, w + h), (synth.uniform_u8(w, h, 3 * w) >= 128).astype(np.uint8) * 255)):
        s = Plane.from_array(src)
        d = Plane(w, h, "U16")
        call(f"CannySobel_U16_U8_{g}x{g}_{'L2' if l2 else 'L1'}NORM", w, h, d, d.stride, s, s.stride)
        got = d.t.cpu().numpy().view(np.uint16)
        assert_same(got, orc.canny_sobel_ex(src, g, l2, padded=True), f"canny sobel noise {k} {g} l2={l2} {w}x{h}")


@pytest.mark.parametrize("w,h", [(100, 60), (333, 77), (640, 480)])
@pytest.mark.parametrize("g,l2", VARIANTS)
def test_canny_node_variants(gpu, orc, w, h, g, l2):
    from mivisionx_amd import openvx as vx
    L = vx.lib()
    ctx = L.vxCreateContext()
    try:
        src = _src(w, h)
        a = vx.create_image(ctx, w, h, "U8")
        vx.write_image(a, src)
        out = vx.create_image(ctx, w, h, "U8")
        lo, hi = THR[g] if not l2 else (THR[g][0] * 3 // 4, THR[g][1] * 3 // 4)
        thr = vx.range_threshold(ctx, lo, hi, "S16")
        graph = L.vxCreateGraph(ctx)
        assert L.vxCannyEdgeDetectorNode(graph, a, thr, g, vx.VX_NORM_L2 if l2 else vx.VX_NORM_L1, out)
        assert L.vxVerifyGraph(graph) == 0
        for rep in range(2):
            assert L.vxProcessGraph(graph) == 0
            if g == 7:
                lo, hi = lo // 4, hi // 4
            assert vx.threshold_range(thr, "S16") == (lo, hi)
            assert_same(vx.read_image(out), orc.canny_ex(src, g, l2, lo, hi), f"canny node {g} l2={l2} {w}x{h} run {rep}")
        gg = ctypes.c_void_p(graph)
        L.vxReleaseGraph(ctypes.byref(gg))
    finally:
        cc = ctypes.c_void_p(ctx)
        L.vxReleaseContext(ctypes.byref(cc))
