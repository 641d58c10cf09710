"""GPU parity through the OpenVX API of libmvx_openvx.so: vx*Node graphs, vxu* immediate calls, the
fused vision supernode versus per-node execution, GDF import, the runvx runner, vxScheduleGraph,
target affinity fallback and zero-copy HIP image handles -- all against the oracle, bit-exact."""
import ctypes
import json
import os
import subprocess

import numpy as np
import pytest

from mivisionx_amd import synth
from util import assert_same, sha

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
c = ctypes


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


def _chain1_graph(vx, ctx, w, h):
    L = vx.lib()
    g = L.vxCreateGraph(ctx)
    src = vx.create_image(ctx, w, h, "U8")
    blur = L.vxCreateVirtualImage(g, 0, 0, vx.DF["VIRT"])
    gx, gy = vx.create_image(ctx, w, h, "S16"), vx.create_image(ctx, w, h, "S16")
    mag, ph = vx.create_image(ctx, w, h, "S16"), vx.create_image(ctx, w, h, "U8")
    assert L.vxGaussian3x3Node(g, src, blur) and L.vxSobel3x3Node(g, blur, gx, gy)
    assert L.vxMagnitudeNode(g, gx, gy, mag) and L.vxPhaseNode(g, gx, gy, ph)
    return g, src, (gx, gy, mag, ph)


@pytest.mark.parametrize("w,h", [(1920, 1080), (333, 77), (10, 10), (5, 3)])
def test_chain1_vx_graph(vx, ctx, orc, w, h):
    """Gaussian (virtual) -> Sobel -> Magnitude + Phase built with vx*Node calls"""
    L = vx.lib()
    g, src, outs = _chain1_graph(vx, ctx, w, h)
    assert L.vxVerifyGraph(g) == 0
    img = synth.uniform_u8(w, h, 0)
    vx.write_image(src, img)
    want = orc.chain1(img)
    for rep in range(2):
        assert L.vxProcessGraph(g) == 0
        for o, wnt, k in zip(outs, want, ("gx", "gy", "mag", "phase")):
            assert_same(vx.read_image(o), wnt, f"vx chain1 {k} {w}x{h} rep {rep}")
    perf = vx.Perf()
    assert L.vxQueryGraph(g, vx.VX_GRAPH_PERFORMANCE, c.byref(perf), c.sizeof(perf)) == 0
    assert perf.num == 2 and perf.avg > 0


def test_chain1_golden_hashes(vx, ctx, golden):
    """the 1920x1080 chain against the reference engine's own output hashes"""
    L = vx.lib()
    w, h = 1920, 1080
    g, src, outs = _chain1_graph(vx, ctx, w, h)
    vx.write_image(src, synth.uniform_u8(w, h, 0))
    assert L.vxProcessGraph(g) == 0    # verifies on first use
    entry = next(e for e in golden["cases"] if (e["w"], e["h"], e["seed"]) == (w, h, 0))
    for o, k in zip(outs, ("gx", "gy", "mag", "phase")):
        assert sha(vx.read_image(o)) == entry["chain1"][k]["sha256"], k


@pytest.mark.parametrize("fn,which", [("vxuGaussian3x3", 0), ("vxuBox3x3", 1), ("vxuMedian3x3", 2)])
def test_vxu_filters(vx, ctx, orc, fn, which):
    L = vx.lib()
    w, h = 97, 41
    img = synth.scene_u8(w, h, 11)
    a, b = vx.create_image(ctx, w, h, "U8"), vx.create_image(ctx, w, h, "U8")
    vx.write_image(a, img)
    assert getattr(L, fn)(ctx, a, b) == 0
    assert_same(vx.read_image(b), orc.filter3x3(which, img), fn)


def test_vxu_sobel_magnitude_phase_scale_pyramid(vx, ctx, orc):
    L = vx.lib()
    w, h = 130, 66
    img = synth.scene_u8(w, h, 12)
    a = vx.create_image(ctx, w, h, "U8")
    vx.write_image(a, img)
    gx, gy = vx.create_image(ctx, w, h, "S16"), vx.create_image(ctx, w, h, "S16")
    assert L.vxuSobel3x3(ctx, a, gx, gy) == 0
    ex, ey = orc.sobel(img)
    assert_same(vx.read_image(gx), ex, "sobel gx")
    m, p = vx.create_image(ctx, w, h, "S16"), vx.create_image(ctx, w, h, "U8")
    assert L.vxuMagnitude(ctx, gx, gy, m) == 0 and L.vxuPhase(ctx, gx, gy, p) == 0
    assert_same(vx.read_image(m), orc.magnitude(ex, ey), "magnitude")
    assert_same(vx.read_image(p), orc.phase(ex, ey), "phase")
    s = vx.create_image(ctx, 77, 50, "U8")
    assert L.vxuScaleImage(ctx, a, s, vx.VX_INTERPOLATION_BILINEAR) == 0
    assert_same(vx.read_image(s), orc.scale(img, 77, 50), "scale")
    assert L.vxuScaleImage(ctx, a, s, vx.VX_INTERPOLATION_NEAREST_NEIGHBOR) == 0
    assert_same(vx.read_image(s), orc.scale_mode(img, 77, 50, 0x4000), "scale nearest")
    pyr = L.vxCreatePyramid(ctx, 3, vx.VX_SCALE_PYRAMID_HALF, w, h, vx.DF["U8"])
    assert L.vxuGaussianPyramid(ctx, a, pyr) == 0
    want = orc.pyramid(img, 3)
    for l in range(3):
        lvl = L.vxGetPyramidLevel(pyr, l)
        assert_same(vx.read_image(lvl), want[l], f"pyramid level {l}")
        vx.release(lvl)


@pytest.mark.parametrize("w,h", [(131, 67), (640, 480)])
def test_vxu_halfscale_gaussian(vx, ctx, orc, w, h):
    """halfscale_gaussian: kernel size 5 is ScaleGaussianHalf 5x5, 3 is ScaleGaussianHalf 3x3, 1 is ScaleImage
    nearest (ago_drama_divide.cpp:1471-1488)"""
    L = vx.lib()
    img = synth.scene_u8(w, h, 13)
    a = vx.create_image(ctx, w, h, "U8")
    vx.write_image(a, img)
    dw, dh = (w + 1) // 2, (h + 1) // 2
    b = vx.create_image(ctx, dw, dh, "U8")
    assert L.vxuHalfScaleGaussian(ctx, a, b, 5) == 0
    assert_same(vx.read_image(b), orc.gaussian_half(img, dw, dh), f"halfscale_gaussian {w}x{h}")
    for k in (1, 3):
        b = vx.create_image(ctx, dw, dh, "U8")   # kernel size 3 leaves rows 0 and dh-1 as they were
        assert L.vxuHalfScaleGaussian(ctx, a, b, k) == 0
        assert_same(vx.read_image(b), orc.halfscale(img, k), f"halfscale_gaussian {k} {w}x{h}")
    assert L.vxuHalfScaleGaussian(ctx, a, b, 7) == vx.VX_ERROR_INVALID_VALUE


def _graph2(vx, ctx, w, h, levels, lower, upper, k, thr, dist, capacity, fuse=True):
    L = vx.lib()
    g = L.vxCreateGraph(ctx)
    if not fuse:
        flags = c.c_uint32(vx.AGO_GRAPH_OPTIMIZER_FLAG_NO_SUPERNODE_MERGE)
        assert L.vxSetGraphAttribute(g, vx.VX_GRAPH_ATTRIBUTE_AMD_OPTIMIZER_FLAGS, c.byref(flags), 4) == 0
    o = dict(rgb=vx.create_image(ctx, w, h, "RGB"))
    yuv = L.vxCreateVirtualImage(g, w, h, vx.DF["IYUV"])
    luma = L.vxCreateVirtualImage(g, 0, 0, vx.DF["VIRT"])
    o["scaled"] = vx.create_image(ctx, w // 2, h // 2, "U8")
    o["pyr"] = L.vxCreatePyramid(ctx, levels, vx.VX_SCALE_PYRAMID_HALF, w, h, vx.DF["U8"])
    o["edges"] = vx.create_image(ctx, w, h, "U8")
    o["hyst"] = vx.range_threshold(ctx, lower, upper)
    o["strength"] = vx.scalar(ctx, vx.VX_TYPE_FLOAT32, thr)
    o["dist"] = vx.scalar(ctx, vx.VX_TYPE_FLOAT32, dist)
    o["k"] = vx.scalar(ctx, vx.VX_TYPE_FLOAT32, k)
    o["corners"] = L.vxCreateArray(ctx, vx.VX_TYPE_KEYPOINT, capacity)
    o["count"] = vx.scalar(ctx, vx.VX_TYPE_SIZE, 0)
    assert L.vxColorConvertNode(g, o["rgb"], yuv)
    assert L.vxChannelExtractNode(g, yuv, vx.VX_CHANNEL_Y, luma)
    assert L.vxScaleImageNode(g, luma, o["scaled"], vx.VX_INTERPOLATION_BILINEAR)
    assert L.vxGaussianPyramidNode(g, luma, o["pyr"])
    assert L.vxCannyEdgeDetectorNode(g, luma, o["hyst"], 3, vx.VX_NORM_L1, o["edges"])
    assert L.vxHarrisCornersNode(g, luma, o["strength"], o["dist"], o["k"], 3, 3, o["corners"], o["count"])
    return g, o


def _graph2_results(vx, o, levels):
    L = vx.lib()
    pyr = []
    for l in range(levels):
        lvl = L.vxGetPyramidLevel(o["pyr"], l)
        pyr.append(vx.read_image(lvl))
        vx.release(lvl)
    return dict(scaled=vx.read_image(o["scaled"]), pyramid=pyr, edges=vx.read_image(o["edges"]),
                corners=vx.read_array(o["corners"], vx.KEYPOINT), count=vx.read_scalar(o["count"], vx.VX_TYPE_SIZE))


@pytest.mark.parametrize("fuse", [True, False])
@pytest.mark.parametrize("w,h", [(644, 482), (100, 64), (38, 22)])
def test_graph2_vx_api(vx, ctx, orc, w, h, fuse):
    """the RGB vision graph through vx*Node calls: fused supernode and per-node execution agree with
    the oracle (and hence with each other)"""
    levels = 5 if min(w, h) >= 80 else 2
    p = dict(lower=80, upper=100, k=0.04, thr=0.0005, dist=5.0, capacity=4096)
    if w * h < 100000:
        p.update(lower=20, upper=60, thr=1e-6)
    L = vx.lib()
    messages = []
    cb = vx.LOG_CB(lambda ctx_, ref, status, text: messages.append(text.decode()))
    L.vxRegisterLogCallback(ctx, cb, 0)
    g, o = _graph2(vx, ctx, w, h, levels, **p, fuse=fuse)
    rgb = synth.scene_rgb(w, h, 50)
    vx.write_image(o["rgb"], rgb)
    assert L.vxVerifyGraph(g) == 0
    # the supernode applies when fusion is on and the frame fits it (width a multiple of 4)
    assert any("fused vision supernode" in m for m in messages) == (fuse and w % 4 == 0), messages
    want = orc.graph2(rgb, w // 2, h // 2, levels, p["lower"], p["upper"], p["k"], p["thr"], p["dist"], p["capacity"])
    for rep in range(2):
        assert L.vxProcessGraph(g) == 0
        got = _graph2_results(vx, o, levels)
        tag = f"{w}x{h} fuse={fuse} rep {rep}"
        assert_same(got["scaled"], want["scaled"], "scaled " + tag)
        for l in range(levels):
            assert_same(got["pyramid"][l], want["pyramid"][l], f"pyramid {l} " + tag)
        assert_same(got["edges"], want["edges"], "edges " + tag)
        assert got["count"] == want["count"], "count " + tag
        assert_same(got["corners"], want["corners"], "corners " + tag)


def test_graph2_4k_vx_api_golden(vx, ctx):
    with open(os.path.join(ROOT, "tests", "golden", "vision_golden_graph2.json")) as f:
        G = json.load(f)["graph2"]
    L = vx.lib()
    g, o = _graph2(vx, ctx, G["w"], G["h"], G["levels"], G["lower"], G["upper"], G["k"], G["thr"], G["dist"], G["capacity"])
    vx.write_image(o["rgb"], synth.scene_rgb(G["w"], G["h"], G["seed"]))
    assert L.vxProcessGraph(g) == 0
    r = _graph2_results(vx, o, G["levels"])
    assert sha(r["scaled"]) == G["scaled"] and [sha(x) for x in r["pyramid"]] == G["pyramid"]
    assert sha(r["edges"]) == G["edges"]
    assert r["count"] == G["count"] and sha(r["corners"]) == G["corners"]


def test_schedule_wait_and_thresholds_read_per_run(vx, ctx, orc):
    """vxScheduleGraph/vxWaitGraph run the graph on the graph's worker; a threshold changed between
    runs is used by the next run (no re-verification needed)"""
    L = vx.lib()
    w, h = 96, 64
    g, o = _graph2(vx, ctx, w, h, 3, 20, 60, 0.04, 1e-6, 5.0, 4096)
    rgb = synth.scene_rgb(w, h, 51)
    vx.write_image(o["rgb"], rgb)
    assert L.vxScheduleGraph(g) == 0
    assert L.vxWaitGraph(g) == 0
    assert_same(vx.read_image(o["edges"]), orc.graph2(rgb, w // 2, h // 2, 3, 20, 60, 0.04, 1e-6, 5.0, 4096)["edges"], "edges")
    lo, hi = vx.PixelValue(), vx.PixelValue()
    lo.U8, hi.U8 = 40, 90
    assert L.vxCopyThresholdRange(o["hyst"], c.byref(lo), c.byref(hi), vx.VX_WRITE_ONLY, vx.VX_MEMORY_TYPE_HOST) == 0
    assert L.vxProcessGraph(g) == 0
    assert_same(vx.read_image(o["edges"]), orc.graph2(rgb, w // 2, h // 2, 3, 40, 90, 0.04, 1e-6, 5.0, 4096)["edges"], "edges 40/90")


def test_cpu_affinity_falls_back_to_gpu(vx, ctx, orc):
    L = vx.lib()
    w, h = 64, 32
    aff = vx.Affinity(vx.AGO_TARGET_AFFINITY_CPU, 0, 0, 0)
    assert L.vxSetContextAttribute(ctx, vx.VX_CONTEXT_ATTRIBUTE_AMD_AFFINITY, c.byref(aff), c.sizeof(aff)) == 0
    img = synth.uniform_u8(w, h, 13)
    a, b = vx.create_image(ctx, w, h, "U8"), vx.create_image(ctx, w, h, "U8")
    vx.write_image(a, img)
    g = L.vxCreateGraph(ctx)
    n = L.vxGaussian3x3Node(g, a, b)
    assert L.vxSetNodeAttribute(n, vx.VX_NODE_ATTRIBUTE_AMD_AFFINITY, c.byref(aff), c.sizeof(aff)) == 0
    assert L.vxProcessGraph(g) == 0
    assert_same(vx.read_image(b), orc.filter3x3(0, img), "gaussian with CPU affinity")


def test_hip_handle_images_are_zero_copy(vx, ctx, orc):
    """vxCreateImageFromHandle(VX_MEMORY_TYPE_HIP) over torch device buffers: the graph reads and
    writes the caller's HBM directly"""
    import torch
    L = vx.lib()
    w, h = 200, 40
    stride = (w + 15) & ~15
    img = synth.scene_u8(w, h, 14)
    src_t = torch.zeros((h, stride), dtype=torch.uint8, device="cuda")
    src_t[:, :w] = torch.from_numpy(img).cuda()
    dst_t = torch.zeros((h, stride), dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    imgs = []
    for t in (src_t, dst_t):
        addr = vx.Addressing(w, h, 1, stride, 1024, 1024, 1, 1, 0)
        ptrs = (c.c_void_p * 1)(t.data_ptr())
        imgs.append(L.vxCreateImageFromHandle(ctx, vx.DF["U8"], c.byref(addr), ptrs, vx.VX_MEMORY_TYPE_HIP))
        assert imgs[-1]
    assert L.vxuMedian3x3(ctx, imgs[0], imgs[1]) == 0
    assert_same(dst_t[:, :w].cpu().numpy(), orc.filter3x3(2, img), "median via HIP handles")


def test_gdf_file_through_runvx(vx, orc, tmp_path):
    """runvx on a GDF: READ the RGB frame, WRITE the outputs, COMPARE against oracle-made files"""
    w, h = 320, 240
    rgb = synth.scene_rgb(w, h, 52)
    want = orc.graph2(rgb, w // 2, h // 2, 4, 80, 100, 0.04, 0.0005, 5.0, 4096)
    (tmp_path / "in.rgb").write_bytes(rgb.tobytes())
    (tmp_path / "want_edges.u8").write_bytes(want["edges"].tobytes())
    gdf = f"""
data rgb = image:{w},{h},RGB2:READ,{tmp_path}/in.rgb
data yuv = image-virtual:{w},{h},IYUV
data luma = image-virtual:{w},{h},U008
data scaled = image:{w // 2},{h // 2},U008:WRITE,{tmp_path}/scaled.u8
data pyr = pyramid:4,HALF,{w},{h},U008
data edges = image:{w},{h},U008:COMPARE,{tmp_path}/want_edges.u8
data hyst = threshold:RANGE,U008,U008:INIT,80,100
data grad = scalar:INT32,3
data strength = scalar:FLOAT32,0.0005
data mind = scalar:FLOAT32,5.0
data sens = scalar:FLOAT32,0.04
data corners = array:KEYPOINT,4096:WRITE,{tmp_path}/corners.bin
data count = scalar:SIZE,0
node org.khronos.openvx.color_convert rgb yuv
node org.khronos.openvx.channel_extract yuv !CHANNEL_Y luma
node org.khronos.openvx.scale_image luma scaled !BILINEAR
node org.khronos.openvx.gaussian_pyramid luma pyr
node org.khronos.openvx.canny_edge_detector luma hyst grad !NORM_L1 edges
node org.khronos.openvx.harris_corners luma strength mind sens grad grad corners count
"""
    (tmp_path / "graph.gdf").write_text(gdf)
    r = subprocess.run([vx.RUNVX, "-frames:3", "-dump-profile", str(tmp_path / "graph.gdf")], capture_output=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout.decode() + r.stderr.decode()
    assert b"3 frame(s)" in r.stdout
    scaled = np.frombuffer((tmp_path / "scaled.u8").read_bytes(), np.uint8).reshape(h // 2, w // 2)
    assert_same(scaled, want["scaled"], "runvx scaled")
    corners = np.frombuffer((tmp_path / "corners.bin").read_bytes(), vx.KEYPOINT)
    assert_same(corners, want["corners"], "runvx corners")
    # a wrong expected file must be reported as a mismatch
    (tmp_path / "want_edges.u8").write_bytes(np.zeros_like(want["edges"]).tobytes())
    r = subprocess.run([vx.RUNVX, str(tmp_path / "graph.gdf")], capture_output=True, timeout=120)
    assert r.returncode == 2 and b"differ" in r.stdout


def test_runvx_node_mode(vx, orc, tmp_path):
    w, h = 128, 72
    img = synth.uniform_u8(w, h, 15)
    (tmp_path / "in.u8").write_bytes(img.tobytes())
    r = subprocess.run([vx.RUNVX, "node", "org.khronos.openvx.gaussian_3x3", f"image:{w},{h},U008:READ,{tmp_path}/in.u8",
                        f"image:{w},{h},U008:WRITE,{tmp_path}/out.u8"], capture_output=True, timeout=120)
    assert r.returncode == 0, r.stderr.decode()
    out = np.frombuffer((tmp_path / "out.u8").read_bytes(), np.uint8).reshape(h, w)
    assert_same(out, orc.filter3x3(0, img), "runvx node gaussian")
