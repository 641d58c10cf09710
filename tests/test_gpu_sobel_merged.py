"""GPU parity of the merged Sobel -> Magnitude / Phase kernels (the reference's node-merge rules,
ago_drama_remove.cpp:155-182, for Sobel gradients that are virtual and read only by Magnitude and/or
Phase): the three C-ABI entry points, the chain-batch supernode with virtual gradients, and graphs built
with vx*Node calls and run through vxProcessGraph -- all bit-exact against the oracle, whose merged
functions are pinned to the reference engine (tests/test_oracle_sobel_merged.py)."""
import ctypes

import numpy as np
import pytest
import torch

from mivisionx_amd import synth
from mivisionx_amd.device import Plane, call
from mivisionx_amd.pipelines import ChainBatch
from util import assert_same

pytestmark = pytest.mark.gpu
c = ctypes

SIZES = [(1920, 1080), (1928, 64), (333, 64), (17, 9), (1920, 31), (100, 3), (5, 4), (8, 8), (4096, 17)]


def _src(w, h, seed):
    return synth.scene_u8(w, h, seed) if (w * h) % 2 else synth.uniform_u8(w, h, seed)


@pytest.mark.parametrize("w,h", SIZES)
def test_abi_merged_kernels_vs_oracle(gpu, orc, w, h):
    print(f"merged sobel ABI {w}x{h}", flush=True)
    img = _src(w, h, 3 + w)
    s = Plane.from_array(img)
    m, p = Plane(w, h, "S16"), Plane(w, h, "U8")
    call("SobelMagnitudePhase_S16U8_U8_3x3", w, h, m, m.stride, p, p.stride, s, s.stride)
    torch.cuda.synchronize()
    wm, wp = orc.sobel_merged(img, 0)
    assert_same(m.download(), wm, f"SobelMagnitudePhase mag {w}x{h}")
    assert_same(p.download(), wp, f"SobelMagnitudePhase phase {w}x{h}")
    m1 = Plane(w, h, "S16")
    call("SobelMagnitude_S16_U8_3x3", w, h, m1, m1.stride, s, s.stride)
    p2 = Plane(w, h, "U8")
    call("SobelPhase_U8_U8_3x3", w, h, p2, p2.stride, s, s.stride)
    torch.cuda.synchronize()
    assert_same(m1.download(), orc.sobel_merged(img, 1)[0], f"SobelMagnitude {w}x{h}")
    assert_same(p2.download(), orc.sobel_merged(img, 2)[1], f"SobelPhase {w}x{h}")


def test_abi_merged_rows_0_and_last_untouched(gpu):
    """the reference wrappers run the merged kernels over rows 1..h-2: rows 0 and h-1 keep their bytes"""
    w, h = 333, 20
    s = Plane.from_array(synth.uniform_u8(w, h, 5))
    m, p, q = Plane(w, h, "S16"), Plane(w, h, "U8"), Plane(w, h, "U8")
    for t in (m.t, p.t, q.t):
        t.fill_(0xA5)
    call("SobelMagnitudePhase_S16U8_U8_3x3", w, h, m, m.stride, p, p.stride, s, s.stride)
    call("SobelPhase_U8_U8_3x3", w, h, q, q.stride, s, s.stride)
    torch.cuda.synchronize()
    for plane, k in ((m, "mag"), (p, "phase"), (q, "phase-only")):
        raw = plane.t.cpu().numpy()
        assert (raw[0] == 0xA5).all() and (raw[-1] == 0xA5).all(), f"{k}: rows 0 / h-1 written"
        assert (raw[1:-1, plane.width * plane.bpp:] == 0xA5).all(), f"{k}: row padding written"


UNEVEN = [(34, 18), (50, 4), (322, 242), (16, 4), (130, 66), (64, 32), (98, 100), (640, 480), (18, 130), (1928, 66)]


@pytest.mark.parametrize("tile", [0, 2, 4])
def test_chain_batch_virtual_gradients_vs_oracle(gpu, orc, tile):
    """one persistent launch with gx = gy = NULL: Gaussian (virtual) -> SobelMagnitudePhase per image"""
    imgs = [_src(w, h, 60 + i) for i, (w, h) in enumerate(UNEVEN)]
    b = ChainBatch([Plane.from_array(a) for a in imgs], virtual_gradients=True).prepare(tile)
    b.run()
    torch.cuda.synchronize()
    for i, a in enumerate(imgs):
        h, w = a.shape
        wm, wp = orc.sobel_merged(a, 0, blur=True)
        got = b.results(i)
        assert_same(got["mag"], wm, f"tile {tile} image {i} {w}x{h} mag")
        assert_same(got["phase"], wp, f"tile {tile} image {i} {w}x{h} phase")
        assert_same(got["half"], orc.scale(a, w // 2, h // 2), f"tile {tile} image {i} half")


# ---------------------------------------------------------------- through the OpenVX engine --------

@pytest.fixture(scope="module")
def vx(gpu):
    from mivisionx_amd import openvx
    return openvx


@pytest.fixture()
def ctx(vx):
    L = vx.lib()
    h = L.vxCreateContext()
    assert h
    yield h
    hh = c.c_void_p(h)
    assert L.vxReleaseContext(c.byref(hh)) == 0


def _merged_graph(vx, ctx, w, h, which, blur, flags=None):
    L = vx.lib()
    g = L.vxCreateGraph(ctx)
    if flags is not None:
        f = c.c_uint32(flags)
        assert L.vxSetGraphAttribute(g, vx.VX_GRAPH_ATTRIBUTE_AMD_OPTIMIZER_FLAGS, c.byref(f), 4) == 0
    src = vx.create_image(ctx, w, h, "U8")
    s = src
    if blur:
        s = L.vxCreateVirtualImage(g, 0, 0, vx.DF["VIRT"])
        assert L.vxGaussian3x3Node(g, src, s)
    gx = L.vxCreateVirtualImage(g, 0, 0, vx.DF["VIRT"])
    gy = L.vxCreateVirtualImage(g, 0, 0, vx.DF["VIRT"])
    assert L.vxSobel3x3Node(g, s, gx, gy)
    mag = vx.create_image(ctx, w, h, "S16") if which != 2 else None
    ph = vx.create_image(ctx, w, h, "U8") if which != 1 else None
    if mag:
        assert L.vxMagnitudeNode(g, gx, gy, mag)
    if ph:
        assert L.vxPhaseNode(g, gx, gy, ph)
    return g, src, mag, ph


@pytest.mark.parametrize("w,h", [(1928, 64), (333, 64), (17, 9), (1920, 1080), (322, 242)])
@pytest.mark.parametrize("which", [0, 1, 2])
@pytest.mark.parametrize("blur", [False, True])
def test_vx_graph_virtual_gradients_merge(vx, ctx, orc, w, h, which, blur):
    """vxProcessGraph of [Gaussian ->] Sobel (virtual gx, gy) -> Magnitude / Phase equals the reference's
    merged kernels; with blur and both outputs at even sizes the chain-batch supernode runs them"""
    print(f"vx merged graph {w}x{h} which={which} blur={blur}", flush=True)
    L = vx.lib()
    g, src, mag, ph = _merged_graph(vx, ctx, w, h, which, blur)
    img = _src(w, h, 17 + w)
    vx.write_image(src, img)
    wm, wp = orc.sobel_merged(img, which, blur)
    for rep in range(2):
        assert L.vxProcessGraph(g) == 0
        if mag:
            assert_same(vx.read_image(mag), wm, f"merged mag {w}x{h} which={which} blur={blur} rep {rep}")
        if ph:
            assert_same(vx.read_image(ph), wp, f"merged phase {w}x{h} which={which} blur={blur} rep {rep}")


@pytest.mark.parametrize("w,h", [(1928, 64), (333, 33)])
def test_vx_graph_no_node_merge_flag_keeps_unmerged_rules(vx, ctx, orc, w, h):
    """AGO_GRAPH_OPTIMIZER_FLAG_NO_NODE_MERGE: the Sobel, Magnitude and Phase kernels run one by one"""
    L = vx.lib()
    img = _src(w, h, 23)
    gx, gy = orc.sobel(img)
    for which in (0, 2):
        g, src, mag, ph = _merged_graph(vx, ctx, w, h, which, False, flags=vx.AGO_GRAPH_OPTIMIZER_FLAG_NO_NODE_MERGE)
        vx.write_image(src, img)
        assert L.vxProcessGraph(g) == 0
        if mag:
            assert_same(vx.read_image(mag), orc.magnitude(gx, gy), f"unmerged mag {w}x{h}")
        assert_same(vx.read_image(ph), orc.phase(gx, gy), f"unmerged phase {w}x{h} which={which}")


def test_vx_graph_gradient_read_elsewhere_is_not_merged(vx, ctx, orc):
    """a third reader of gx (a second Magnitude node) keeps every node unmerged (the SOLITARY rule)"""
    L = vx.lib()
    w, h = 1928, 40
    img = _src(w, h, 29)
    g = L.vxCreateGraph(ctx)
    src = vx.create_image(ctx, w, h, "U8")
    gx = L.vxCreateVirtualImage(g, 0, 0, vx.DF["VIRT"])
    gy = L.vxCreateVirtualImage(g, 0, 0, vx.DF["VIRT"])
    m1, m2, p = vx.create_image(ctx, w, h, "S16"), vx.create_image(ctx, w, h, "S16"), vx.create_image(ctx, w, h, "U8")
    assert L.vxSobel3x3Node(g, src, gx, gy) and L.vxMagnitudeNode(g, gx, gy, m1) and L.vxPhaseNode(g, gx, gy, p)
    assert L.vxMagnitudeNode(g, gy, gx, m2)
    vx.write_image(src, img)
    assert L.vxProcessGraph(g) == 0
    ex, ey = orc.sobel(img)
    assert_same(vx.read_image(m1), orc.magnitude(ex, ey), "mag (gx, gy)")
    assert_same(vx.read_image(m2), orc.magnitude(ey, ex), "mag (gy, gx)")
    assert_same(vx.read_image(p), orc.phase(ex, ey), "phase")


def test_gdf_amd_merged_kernel_names(vx, ctx, orc):
    """the AMD merged kernels named directly in a GDF (com.amd.openvx.SobelMagnitudePhase_S16U8_U8_3x3 ...)"""
    L = vx.lib()
    w, h = 333, 50
    img = _src(w, h, 31)
    g = L.vxCreateGraph(ctx)
    src = vx.create_image(ctx, w, h, "U8")
    m, p, m1, p1 = (vx.create_image(ctx, w, h, f) for f in ("S16", "U8", "S16", "U8"))
    # $n names the n-th reference handed to the import (the GDF grammar has no "data x = $n" form)
    text = ("node com.amd.openvx.SobelMagnitudePhase_S16U8_U8_3x3 $2 $3 $1\n"
            "node com.amd.openvx.SobelMagnitude_S16_U8_3x3 $4 $1\n"
            "node com.amd.openvx.SobelPhase_U8_U8_3x3 $5 $1\n")
    assert vx.import_gdf(g, text, refs=(src, m, p, m1, p1)) == 0
    vx.write_image(src, img)
    assert L.vxProcessGraph(g) == 0
    wm, wp = orc.sobel_merged(img, 0)
    assert_same(vx.read_image(m), wm, "gdf SobelMagnitudePhase mag")
    assert_same(vx.read_image(p), wp, "gdf SobelMagnitudePhase phase")
    assert_same(vx.read_image(m1), orc.sobel_merged(img, 1)[0], "gdf SobelMagnitude")
    assert_same(vx.read_image(p1), orc.sobel_merged(img, 2)[1], "gdf SobelPhase")
