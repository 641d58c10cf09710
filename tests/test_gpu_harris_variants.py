"""GPU parity of Harris with gradient and block sizes 5 and 7: the HG3 and score planes through the C ABI
(raw plane floats, so the reference's packed writer rows and its reader's row pitch are compared as laid
out) and the whole harris_corners node through the OpenVX API -- against the oracle, bit-exact."""
import ctypes

import numpy as np
import pytest
import torch

from mivisionx_amd import synth
from mivisionx_amd.device import Plane, call
from util import assert_same

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("w,h", [(64, 48), (101, 61), (37, 29), (333, 77), (1920, 1080)])
@pytest.mark.parametrize("g,b", [(5, 3), (7, 3), (3, 5), (3, 7), (5, 5), (7, 7), (5, 7)])
def test_harris_planes(gpu, orc, w, h, g, b):
    src = synth.scene_u8(w, h, 11)
    s = Plane.from_array(src)
    hs, vs = (12 * w + 15) & ~15, (4 * w + 15) & ~15
    hg3 = torch.zeros((h, hs // 4), dtype=torch.float32, device="cuda")
    vc = torch.zeros((h, vs // 4), dtype=torch.float32, device="cuda")
    hp, vp = ctypes.c_void_p(hg3.data_ptr()), ctypes.c_void_p(vc.data_ptr())
    call(f"HarrisSobel_HG3_U8_{g}x{g}", w, h, hp, hs, s, s.stride)
    norm = float(orc.lib.oracle_harris_norm_gb(g, b))
    call(f"HarrisScore_HVC_HG3_{b}x{b}", w, h, vp, vs, hp, hs, 0.04, 1e-5, norm)
    want_hg3, want_vc = orc.harris_planes(src, g, b, 0.04, 1e-5)
    assert_same(hg3.cpu().numpy().view(np.uint32), want_hg3.view(np.uint32), f"HG3 {g} {w}x{h}")
    assert_same(vc.cpu().numpy()[:, :w].view(np.uint32), want_vc.view(np.uint32), f"HVC {g}/{b} {w}x{h}")


@pytest.mark.parametrize("w,h", [(101, 61), (640, 480)])
@pytest.mark.parametrize("g,b", [(5, 5), (7, 7), (3, 7), (7, 3)])
def test_harris_node_variants(gpu, orc, w, h, g, b):
    from mivisionx_amd import openvx as vx
    L = vx.lib()
    ctx = L.vxCreateContext()
    try:
        src = synth.scene_u8(w, h, 12)
        a = vx.create_image(ctx, w, h, "U8")
        vx.write_image(a, src)
        cap = 3000
        arr = L.vxCreateArray(ctx, vx.VX_TYPE_KEYPOINT, cap)
        num = L.vxCreateScalar(ctx, vx.VX_TYPE_SIZE, ctypes.byref(ctypes.c_size_t(0)))
        f = lambda v: L.vxCreateScalar(ctx, vx.VX_TYPE_FLOAT32, ctypes.byref(ctypes.c_float(v)))
        graph = L.vxCreateGraph(ctx)
        assert L.vxHarrisCornersNode(graph, a, f(1e-5), f(3.0), f(0.04), g, b, arr, num)
        assert L.vxProcessGraph(graph) == 0
        got = vx.read_array(arr, vx.KEYPOINT)
        want, count = orc.harris_ex(src, g, b, 0.04, 1e-5, 3.0, cap)
        assert vx.read_scalar(num, vx.VX_TYPE_SIZE) == count
        assert len(got) == len(want) and np.array_equal(got.view(np.uint8).ravel(), want.view(np.uint8).ravel())
        gg = ctypes.c_void_p(graph)
        L.vxReleaseGraph(ctypes.byref(gg))
    finally:
        cc = ctypes.c_void_p(ctx)
        L.vxReleaseContext(ctypes.byref(cc))
