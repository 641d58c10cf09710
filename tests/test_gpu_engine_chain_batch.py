"""The engine's chain-batch supernode through the OpenVX API: a graph holding many
Gaussian3x3 -> Sobel3x3 -> Magnitude + Phase chains (and half-size ScaleImage nodes) over images of
uneven sizes runs them as one batched launch replayed from a HIP graph; outputs are bit-exact with
the oracle and with per-node execution (optimizer merge off), chains that do not qualify (odd sizes,
inputs produced inside the graph) run per node beside the batch, and the AMD low-level kernel names
of a GDF graph are batched the same way."""
import ctypes
import os

import numpy as np
import pytest

from mivisionx_amd import synth
from util import assert_same

pytestmark = pytest.mark.gpu

c = ctypes
SIZES = [(130, 66), (64, 32), (322, 242), (34, 18), (640, 480), (98, 100), (1920, 1080)]


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


def _oracle(orc, img, with_half=True):
    h, w = img.shape
    gx, gy, mag, ph = orc.chain1(img)
    out = dict(gx=gx, gy=gy, mag=mag, phase=ph)
    if with_half:
        out["half"] = orc.scale(img, w // 2, h // 2)
    return out


def _chain(vx, ctx, g, w, h, src=None, half=True):
    L = vx.lib()
    src = src or vx.create_image(ctx, w, h, "U8")
    blur = L.vxCreateVirtualImage(g, 0, 0, vx.DF["VIRT"])
    o = dict(src=src, gx=vx.create_image(ctx, w, h, "S16"), gy=vx.create_image(ctx, w, h, "S16"),
             mag=vx.create_image(ctx, w, h, "S16"), phase=vx.create_image(ctx, w, h, "U8"))
    assert L.vxGaussian3x3Node(g, src, blur) and L.vxSobel3x3Node(g, blur, o["gx"], o["gy"])
    assert L.vxMagnitudeNode(g, o["gx"], o["gy"], o["mag"]) and L.vxPhaseNode(g, o["gx"], o["gy"], o["phase"])
    if half:
        o["half"] = vx.create_image(ctx, w // 2, h // 2, "U8")
        assert L.vxScaleImageNode(g, src, o["half"], vx.VX_INTERPOLATION_BILINEAR)
    return o


@pytest.mark.parametrize("merge", [True, False])
def test_many_chains_one_batch(vx, ctx, orc, merge):
    L = vx.lib()
    messages = []
    cb = vx.LOG_CB(lambda ctx_, ref, status, text: messages.append(text.decode()))
    L.vxRegisterLogCallback(ctx, cb, 0)
    g = L.vxCreateGraph(ctx)
    if not merge:
        flags = c.c_uint32(vx.AGO_GRAPH_OPTIMIZER_FLAG_NO_SUPERNODE_MERGE)
        assert L.vxSetGraphAttribute(g, vx.VX_GRAPH_ATTRIBUTE_AMD_OPTIMIZER_FLAGS, c.byref(flags), 4) == 0
    chains = [_chain(vx, ctx, g, w, h, half=(i % 3 != 1)) for i, (w, h) in enumerate(SIZES)]
    # not batched: an odd size, and a chain whose input another node of the graph writes
    odd = _chain(vx, ctx, g, 333, 77, half=False)
    boxed = vx.create_image(ctx, 96, 40, "U8")
    boxsrc = vx.create_image(ctx, 96, 40, "U8")
    assert L.vxBox3x3Node(g, boxsrc, boxed)
    inner = _chain(vx, ctx, g, 96, 40, src=boxed, half=False)
    assert L.vxVerifyGraph(g) == 0
    batched = [m for m in messages if "chain batch supernode" in m]
    if merge:
        assert len(batched) == 1 and f"{len(SIZES)} chains" in batched[0], messages
    else:
        assert not batched, messages
    imgs = [synth.scene_u8(w, h, 70 + i) for i, (w, h) in enumerate(SIZES)]
    odd_img = synth.uniform_u8(333, 77, 3)
    box_img = synth.uniform_u8(96, 40, 4)
    for rep in range(2):
        if rep:   # new inputs between runs: the replayed launch reads them
            imgs = [synth.uniform_u8(w, h, 90 + i) for i, (w, h) in enumerate(SIZES)]
        for o, img in zip(chains, imgs):
            vx.write_image(o["src"], img)
        vx.write_image(odd["src"], odd_img)
        vx.write_image(boxsrc, box_img)
        assert L.vxProcessGraph(g) == 0
        for i, (o, img) in enumerate(zip(chains, imgs)):
            want = _oracle(orc, img, "half" in o)
            for k, v in want.items():
                assert_same(vx.read_image(o[k]), v, f"chain {i} {SIZES[i]} {k} merge={merge} rep {rep}")
        for k, v in _oracle(orc, odd_img, False).items():
            assert_same(vx.read_image(odd[k]), v, f"odd-size chain {k}")
        for k, v in _oracle(orc, orc.filter3x3(1, box_img), False).items():
            assert_same(vx.read_image(inner[k]), v, f"in-graph input chain {k}")


def test_gdf_amd_kernel_names_batched(vx, ctx, orc):
    """the AMD low-level names (output first) of the same chains, imported from GDF text"""
    L = vx.lib()
    messages = []
    cb = vx.LOG_CB(lambda ctx_, ref, status, text: messages.append(text.decode()))
    L.vxRegisterLogCallback(ctx, cb, 0)
    g = L.vxCreateGraph(ctx)
    sizes = [(256, 128), (98, 64)]
    refs, lines = [], []
    for i, (w, h) in enumerate(sizes):
        imgs = [vx.create_image(ctx, w, h, f) for f in ("U8", "S16", "S16", "S16", "U8")] + \
               [vx.create_image(ctx, w // 2, h // 2, "U8")]
        b = len(refs)
        refs += imgs
        lines += [f"data blur{i} = virtual-image:{w},{h},U008",
                  f"node com.amd.openvx.Gaussian_U8_U8_3x3 blur{i} ${b + 1}",
                  f"node com.amd.openvx.Sobel_S16S16_U8_3x3_GXY ${b + 2} ${b + 3} blur{i}",
                  f"node com.amd.openvx.Magnitude_S16_S16S16 ${b + 4} ${b + 2} ${b + 3}",
                  f"node com.amd.openvx.Phase_U8_S16S16 ${b + 5} ${b + 2} ${b + 3}",
                  f"node com.amd.openvx.ScaleImage_U8_U8_Bilinear ${b + 6} ${b + 1}"]
    assert vx.import_gdf(g, "\n".join(lines) + "\n", refs) == 0
    assert L.vxVerifyGraph(g) == 0
    assert any("chain batch supernode: 2 chains" in m for m in messages), messages
    inputs = [synth.scene_u8(w, h, 80 + i) for i, (w, h) in enumerate(sizes)]
    for i, img in enumerate(inputs):
        vx.write_image(refs[6 * i], img)
    assert L.vxProcessGraph(g) == 0
    for i, img in enumerate(inputs):
        want = _oracle(orc, img)
        for j, k in enumerate(("gx", "gy", "mag", "phase", "half")):
            assert_same(vx.read_image(refs[6 * i + 1 + j]), want[k], f"gdf chain {i} {k}")
