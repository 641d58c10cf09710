"""GPU parity of FastCorners (with and without non-maximum suppression): the C ABI and the fast_corners
node / vxuFastCorners through the OpenVX API, against the oracle -- keypoint arrays bit-exact, counts
exact, capacity truncation in raster order."""
import ctypes

import numpy as np
import pytest
import torch

from mivisionx_amd import synth
from mivisionx_amd.device import Plane, call
from pyoracle import KEYPOINT

pytestmark = pytest.mark.gpu


def _run(src, thr, nonmax, capacity):
    h, w = src.shape
    s = Plane.from_array(src)
    corners = torch.zeros(max(capacity, 1) * 28, dtype=torch.uint8, device="cuda")
    cnt = torch.zeros(1, dtype=torch.int32, device="cuda")
    num = torch.zeros(1, dtype=torch.int64, device="cuda")
    from mivisionx_amd import kernels
    K = kernels()
    K.mvx_FastCorners_workspace_size.restype = ctypes.c_size_t
    wsz = K.mvx_FastCorners_workspace_size(w, h)
    ws = torch.zeros(int(wsz), dtype=torch.uint8, device="cuda")
    vp = lambda t: ctypes.c_void_p(t.data_ptr())
    call("FastCorners_XY_U8_Supression" if nonmax else "FastCorners_XY_U8_NoSupression", capacity, vp(corners), vp(cnt), vp(num),
         w, h, s, s.stride, float(thr), vp(ws), ctypes.c_size_t(int(wsz)))
    n = int(cnt.item())
    kp = corners.cpu().numpy()[:n * 28].view(KEYPOINT)
    return kp, n, int(num.item())


@pytest.mark.parametrize("w,h", [(64, 48), (101, 61), (37, 29), (333, 77), (8, 8), (1920, 1080)])
@pytest.mark.parametrize("thr", [0.0, 10.0, 25.5, 60.0])
@pytest.mark.parametrize("nonmax", [0, 1])
def test_fast_c_abi(gpu, orc, w, h, thr, nonmax):
    src = synth.scene_u8(w, h, 21) if w >= 16 and h >= 16 else synth.uniform_u8(w, h, 21)
    cap = 200000 if w * h > 1_000_000 else 5000
    kp, n, total = _run(src, thr, nonmax, cap)
    want, wc = orc.fast(src, thr, nonmax, cap)
    assert total == wc and n == len(want)
    assert np.array_equal(kp.view(np.uint8), want.view(np.uint8))


@pytest.mark.parametrize("thr", [200.0, 250.0, 253.0, 254.0, 255.0, -3.5, 0.75])
@pytest.mark.parametrize("nonmax", [0, 1])
def test_fast_threshold_edges(gpu, orc, thr, nonmax):
    """the strength closed form at its edges (R-1 against 254, t >= 255, negative and fractional
    thresholds through (short)floorf) on a high-contrast input whose nine-runs pass up to 254"""
    rng = np.random.default_rng(8)
    for w, h in ((97, 53), (1920, 1080)):
        src = rng.choice(np.array([0, 1, 3, 252, 254, 255], np.uint8), size=(h, w))
        cap = 2_000_000 if w * h > 1_000_000 else 5000
        kp, n, total = _run(src, thr, nonmax, cap)
        want, wc = orc.fast(src, thr, nonmax, cap)
        assert total == wc and n == len(want)
        assert np.array_equal(kp.view(np.uint8), want.view(np.uint8))


def test_fast_capacity(gpu, orc):
    src = synth.scene_u8(333, 77, 4)
    kp, n, total = _run(src, 12.0, 0, 50)
    want, wc = orc.fast(src, 12.0, 0, 50)
    assert n == 50 and total == wc > 50 and np.array_equal(kp.view(np.uint8), want.view(np.uint8))


@pytest.mark.parametrize("nonmax", [0, 1])
def test_fast_node(gpu, orc, nonmax):
    from mivisionx_amd import openvx as vx
    L = vx.lib()
    ctx = L.vxCreateContext()
    try:
        w, h = 640, 480
        src = synth.scene_u8(w, h, 5)
        a = vx.create_image(ctx, w, h, "U8")
        vx.write_image(a, src)
        arr = L.vxCreateArray(ctx, vx.VX_TYPE_KEYPOINT, 4000)
        num = L.vxCreateScalar(ctx, vx.VX_TYPE_SIZE, ctypes.byref(ctypes.c_size_t(0)))
        thr = L.vxCreateScalar(ctx, vx.VX_TYPE_FLOAT32, ctypes.byref(ctypes.c_float(20.0)))
        graph = L.vxCreateGraph(ctx)
        assert L.vxFastCornersNode(graph, a, thr, 1 if nonmax else 0, arr, num)
        for _ in range(2):
            assert L.vxProcessGraph(graph) == 0
            want, wc = orc.fast(src, 20.0, nonmax, 4000)
            assert vx.read_scalar(num, vx.VX_TYPE_SIZE) == wc
            got = vx.read_array(arr, vx.KEYPOINT)
            assert np.array_equal(got.view(np.uint8).ravel(), want.view(np.uint8).ravel())
        arr2 = L.vxCreateArray(ctx, vx.VX_TYPE_KEYPOINT, 4000)
        assert L.vxuFastCorners(ctx, a, thr, 1 if nonmax else 0, arr2, None) == 0
        assert np.array_equal(vx.read_array(arr2, vx.KEYPOINT).view(np.uint8).ravel(), want.view(np.uint8).ravel())
        gg = ctypes.c_void_p(graph)
        L.vxReleaseGraph(ctypes.byref(gg))
    finally:
        cc = ctypes.c_void_p(ctx)
        L.vxReleaseContext(ctypes.byref(cc))


@pytest.mark.parametrize("nonmax", [0, 1])
def test_fast_unaligned_source(gpu, orc, nonmax):
    """a source at an odd address with an odd stride: the tile staging's byte path (aligned planes take
    dword loads)"""
    from mivisionx_amd import kernels
    w, h, stride = 333, 77, 341
    src = synth.scene_u8(w, h, 22)
    buf = torch.zeros(h * stride + 8, dtype=torch.uint8)
    view = buf[1:1 + h * stride].view(h, stride)
    view[:, :w] = torch.from_numpy(src)
    dbuf = buf.cuda()
    corners = torch.zeros(5000 * 28, dtype=torch.uint8, device="cuda")
    cnt = torch.zeros(1, dtype=torch.int32, device="cuda")
    num = torch.zeros(1, dtype=torch.int64, device="cuda")
    K = kernels()
    K.mvx_FastCorners_workspace_size.restype = ctypes.c_size_t
    wsz = K.mvx_FastCorners_workspace_size(w, h)
    ws = torch.zeros(int(wsz), dtype=torch.uint8, device="cuda")
    vp = lambda t: ctypes.c_void_p(t.data_ptr())
    call("FastCorners_XY_U8_Supression" if nonmax else "FastCorners_XY_U8_NoSupression", 5000, vp(corners), vp(cnt),
         vp(num), w, h, ctypes.c_void_p(dbuf.data_ptr() + 1), stride, 20.0, vp(ws), ctypes.c_size_t(int(wsz)))
    n = int(cnt.item())
    kp = corners.cpu().numpy()[:n * 28].view(KEYPOINT)
    want, wc = orc.fast(src, 20.0, nonmax, 5000)
    assert int(num.item()) == wc and n == len(want)
    assert np.array_equal(kp.view(np.uint8), want.view(np.uint8))
