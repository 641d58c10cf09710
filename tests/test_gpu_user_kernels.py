"""GPU run of the plug-in surface: a user-kernel module (tests/hip/user_module.hip) loaded with
vxLoadKernels, whose kernels run in a graph beside hot-path nodes --
  Gaussian3x3 -> invert (GPU user kernel: HIP launch on VX_NODE_ATTRIBUTE_AMD_HIP_STREAM over
  VX_IMAGE_ATTRIBUTE_AMD_HIP_BUFFER pointers) -> Box3x3, add_sat (CPU user kernel through the map API,
  local data and an output scalar), copy_gpubuf (CPU kernel with GPU buffer access)
-- checked against the oracle's filters and numpy; the valid rectangles the graph propagates; and the same
module through runvx, a GDF with an `import` line, in a child process."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from mivisionx_amd import synth
from util import assert_same

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MODULE_DIR = os.path.join(ROOT, "tests", "hip")
c = ctypes


@pytest.fixture(scope="module")
def vx(gpu):
    from mivisionx_amd import openvx
    L = openvx.lib()
    c.CDLL(os.path.join(MODULE_DIR, "libmvx_test_user.so"))   # found by its soname by vxLoadKernels
    L.vxLoadKernels.argtypes = [c.c_void_p, c.c_char_p]
    L.vxGetValidRegionImage.argtypes = [c.c_void_p, c.POINTER(openvx.Rect)]
    return openvx


def _rect(vx, img):
    r = vx.Rect()
    assert vx.lib().vxGetValidRegionImage(img, c.byref(r)) == 0
    return (r.start_x, r.start_y, r.end_x, r.end_y)


def _node(L, ctx, g, name, *params):
    n = L.vxCreateGenericNode(g, L.vxGetKernelByName(ctx, name))
    assert n, name
    for i, p in enumerate(params):
        assert L.vxSetParameterByIndex(n, i, p) == 0
    return n


def test_user_kernels_beside_hot_path_nodes(vx, orc):
    L = vx.lib()
    ctx = L.vxCreateContext()
    try:
        assert L.vxLoadKernels(ctx, b"mvx_test_user") == 0
        w, h = 333, 120
        img = synth.scene_u8(w, h, 77)
        g = L.vxCreateGraph(ctx)
        src = vx.create_image(ctx, w, h, "U8")
        vx.write_image(src, img)
        blur = L.vxCreateVirtualImage(g, 0, 0, vx.DF["VIRT"])
        inv = vx.create_image(ctx, w, h, "U8")
        boxed, total, copied = (vx.create_image(ctx, w, h, "U8") for _ in range(3))
        runs = vx.scalar(ctx, vx.VX_TYPE_UINT32, 0)
        assert L.vxGaussian3x3Node(g, src, blur)
        _node(L, ctx, g, b"com.example.invert", blur, inv)
        assert L.vxBox3x3Node(g, inv, boxed)
        _node(L, ctx, g, b"com.example.add_sat", src, inv, total, runs)
        _node(L, ctx, g, b"com.example.copy_gpubuf", boxed, copied)
        assert L.vxVerifyGraph(g) == 0
        want_inv = (255 - orc.filter3x3(0, img)).astype(np.uint8)
        want_box = orc.filter3x3(1, want_inv)
        want_sum = np.minimum(img.astype(np.int32) + want_inv, 255).astype(np.uint8)
        for rep in range(3):
            assert L.vxProcessGraph(g) == 0
            assert_same(vx.read_image(inv), want_inv, f"invert (GPU user kernel) rep {rep}")
            assert_same(vx.read_image(boxed), want_box, f"box after the user kernel rep {rep}")
            assert_same(vx.read_image(total), want_sum, f"add_sat (CPU user kernel) rep {rep}")
            assert_same(vx.read_image(copied), want_box, f"copy_gpubuf (GPU buffer access) rep {rep}")
            assert vx.read_scalar(runs, vx.VX_TYPE_UINT32) == rep + 1
        # valid rectangles: Gaussian shrinks by 1, the module's callback by 2 more, Box by 1 more; the CPU
        # kernel without a callback resets its output to the whole image
        assert _rect(vx, inv) == (3, 3, w - 3, h - 3)
        assert _rect(vx, boxed) == (4, 4, w - 4, h - 4)
        assert _rect(vx, total) == (0, 0, w, h)
    finally:
        cc = c.c_void_p(ctx)
        L.vxReleaseContext(c.byref(cc))


def test_builtin_valid_rectangles(vx):
    """valid_rect_callback rules of the node set: 3x3 filters and Sobel shrink by 1, Magnitude
    intersects, Canny 3x3 shrinks by 2, pyramid levels halve, warps leave the output whole"""
    L = vx.lib()
    ctx = L.vxCreateContext()
    try:
        w, h = 128, 96
        g = L.vxCreateGraph(ctx)
        src = vx.create_image(ctx, w, h, "U8")
        vx.write_image(src, synth.uniform_u8(w, h, 3))
        gauss = vx.create_image(ctx, w, h, "U8")
        gx, gy, mag = (vx.create_image(ctx, w, h, "S16") for _ in range(3))
        edges = vx.create_image(ctx, w, h, "U8")
        pyr = L.vxCreatePyramid(ctx, 3, vx.VX_SCALE_PYRAMID_HALF, w, h, vx.DF["U8"])
        assert L.vxGaussian3x3Node(g, src, gauss) and L.vxSobel3x3Node(g, gauss, gx, gy) and L.vxMagnitudeNode(g, gx, gy, mag)
        assert L.vxCannyEdgeDetectorNode(g, src, vx.range_threshold(ctx, 80, 100), 3, vx.VX_NORM_L1, edges)
        assert L.vxGaussianPyramidNode(g, gauss, pyr)
        assert L.vxVerifyGraph(g) == 0
        assert _rect(vx, gauss) == (1, 1, w - 1, h - 1)
        assert _rect(vx, gx) == (2, 2, w - 2, h - 2) and _rect(vx, mag) == (2, 2, w - 2, h - 2)
        assert _rect(vx, edges) == (2, 2, w - 2, h - 2)
        lv = [L.vxGetPyramidLevel(pyr, i) for i in range(3)]
        assert _rect(vx, lv[0]) == (1, 1, w - 1, h - 1)
        assert _rect(vx, lv[1]) == (1, 1, w // 2, h // 2) and _rect(vx, lv[2]) == (1, 1, w // 4, h // 4)
        assert _rect(vx, src) == (0, 0, w, h)
    finally:
        cc = c.c_void_p(ctx)
        L.vxReleaseContext(c.byref(cc))


def test_runvx_imports_the_module(vx, orc, tmp_path):
    """runvx with `import mvx_test_user` in the GDF, the module found through LD_LIBRARY_PATH"""
    w, h = 200, 64
    img = synth.uniform_u8(w, h, 5)
    img.tofile(tmp_path / "in.u8")
    gdf = (f"import mvx_test_user\n"
           f"data in = image:{w},{h},U008:READ,in.u8\n"
           f"data g = image-virtual:{w},{h},U008\n"
           f"data inv = image:{w},{h},U008:WRITE,inv.u8\n"
           f"data sum = image:{w},{h},U008:WRITE,sum.u8\n"
           f"data runs = scalar:UINT32,0\n"
           f"node org.khronos.openvx.gaussian_3x3 in g\n"
           f"node com.example.invert g inv\n"
           f"node com.example.add_sat in inv sum runs\n")
    (tmp_path / "user.gdf").write_text(gdf)
    env = dict(os.environ, LD_LIBRARY_PATH=MODULE_DIR + ":" + os.environ.get("LD_LIBRARY_PATH", ""))
    print("runvx with the user module", flush=True)
    r = subprocess.run([vx.RUNVX, "-frames:2", str(tmp_path / "user.gdf")], cwd=tmp_path, env=env, capture_output=True, timeout=120)
    assert r.returncode == 0, r.stdout.decode() + r.stderr.decode()
    want_inv = (255 - orc.filter3x3(0, img)).astype(np.uint8)
    assert_same(np.fromfile(tmp_path / "inv.u8", np.uint8).reshape(-1, h, w)[-1], want_inv, "runvx invert")
    want_sum = np.minimum(img.astype(np.int32) + want_inv, 255).astype(np.uint8)
    assert_same(np.fromfile(tmp_path / "sum.u8", np.uint8).reshape(-1, h, w)[-1], want_sum, "runvx add_sat")
