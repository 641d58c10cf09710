"""Benchmark: the 3840x2160 vision graph on one MI355X.

Workload (one "step" = one pass of the graph over a batch of B resident 4K RGB frames):
  RGB 3840x2160 -> ColorConvert IYUV (virtual) -> Y (virtual);
  Y -> ScaleImage 1920x1080 bilinear; Y -> GaussianPyramid 5 levels HALF;
  Y -> CannyEdgeDetector (3, L1, 80/100); Y -> HarrisCorners (k 0.04, T 0.0005, d 5, 3, 3).
Inputs are synthetic scenes (sum of sinusoids + rectangles + noise, mivisionx_amd.synth), already in
HBM when the timed region starts.  The graph runs through mvx_VisionGraph_run_stage_range (fused tile
pass, hysteresis union, hysteresis resolve, pyramid levels 2-4, Harris pick) in the schedule
mvx_VisionGraph_run uses: the tile pass of all frames and then the hysteresis stages on the main
stream, the pyramid levels and pick on a second stream beside the hysteresis (--chunk N: the frames in
chunks of N, a chunk's later stages on side streams overlapping the next chunk's tile pass).
value = whole-job graph throughput in Mpix/s of 4K input.

Also reported:
  roofline       -- the dominant stage's algorithmic bytes per launch / its mean launch time, timed with
                    HIP events on its launch stream inside the timed region;
  nodes          -- the other node workloads, each timed alone with HIP events on its launch stream:
                    the 1920x1080 Gaussian->Sobel->Magnitude+Phase chain (8 B/px), WarpAffine and
                    WarpPerspective 3840x2160 bilinear constant-0 (2 B/px), Remap 3840x2160 through the
                    fixed-point table of a barrel map (6 B/px), OpticalFlowPyrLK 4 levels / 9x9 / 20
                    iterations / 10,000 keypoints on 3840x2160 (keypoints/s);
  graph_hbm      -- the whole graph's algorithmic bytes per frame / frame time;
  batched_uneven -- 256 U8 images of uneven sizes (W 320..4096, H 240..2160), each Gaussian3x3 ->
                    Sobel3x3 -> Magnitude + Phase plus ScaleImage to half size, through the OpenVX
                    engine: the chain-batch supernode (one persistent launch from a HIP graph) versus
                    per-node launches (merge off), and the batch kernel's achieved GB/s (8.25 B/px);
  alignment      -- 64 pairs of 1024x1024 F32 images, Adam on the 3-level pyramid L2 loss through the
                    fused warp + loss + adjoint kernel: time per step and recovered homography error;
  per_image      -- the same frames through one launch per node and frame (the reference's execution
                    model), and the batched/per-image ratio;
  cpu_baseline   -- the reference's own CPU engine (oracle/_ref, one thread) on a bounded sample of the
                    same workload, or the oracle's C port when oracle/_ref is absent.
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

HBM_PEAK_GBS = 8000.0
W, H, DW, DH, LEVELS = 3840, 2160, 1920, 1080, 5
PARAMS = dict(lower=80, upper=100, k=0.04, thr=0.0005, dist=5.0, capacity=65536)
STAGES = ["fused_tile_pass", "hysteresis_border_union", "hysteresis_resolve", "pyramid_levels_2_4", "harris_pick"]


def pyramid_bytes(w, h, levels):
    tot = 0
    for _ in range(levels):
        tot += w * h
        w, h = (w + 1) // 2, (h + 1) // 2
    return tot


def graph_algorithmic_bytes():
    """graph inputs read once + non-virtual outputs written once (virtual intermediates count 0)"""
    return 3 * W * H + DW * DH + pyramid_bytes(W, H, LEVELS) + W * H


def stage_algorithmic_bytes(stage):
    """graph input/output bytes one frame of each stage reads or writes"""
    px = W * H
    l1 = ((W + 1) // 2) * ((H + 1) // 2)
    return {
        0: 3 * px + px + DW * DH + l1 + px,      # RGB in; level 0, scaled, level 1, edges out
        1: 0,                                     # tile-border tables only
        2: 0,                                     # re-reads/writes edge pixels already counted
        3: pyramid_bytes(W, H, LEVELS) - px,      # level 1 in; levels 2.. out
        4: 0,
    }[stage]


def measure_nodes(reps=20):
    """GPU time of the node workloads beside the headline graph (HIP events on the launch stream)"""
    import ctypes
    import numpy as np
    import torch
    from mivisionx_amd import synth
    from mivisionx_amd.device import Plane, call
    from mivisionx_amd.pipelines import PyrLKTracker, pyramid_dims

    def timed(fn):
        """mean device time of one call over `reps` back-to-back calls bracketed by one pair of HIP
        events on the launch stream, after enough warm-up calls (about 2 ms of work) that the GPU
        clocks have ramped up"""
        t0 = time.perf_counter()
        for _ in range(100):
            fn()
            torch.cuda.synchronize()
            if time.perf_counter() - t0 > 2e-3:
                break
        for _ in range(5):
            fn()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(reps):
            fn()
        b.record()
        torch.cuda.synchronize()
        return a.elapsed_time(b) / reps

    def bw(name, ms, px, bpp, unit_value, unit):
        gbs = px * bpp / (ms / 1e3) / 1e9
        return {"ms": round(ms, 4), "value": round(unit_value, 2), "unit": unit, "algorithmic_bytes": int(px * bpp),
                "achieved_GBps": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4), "bound": "hbm"}

    out = {}
    # the 1920x1080 Gaussian -> Sobel -> Magnitude + Phase chain, one launch per node
    w1, h1 = 1920, 1080
    s = Plane.from_array(synth.uniform_u8(w1, h1, 0))
    g, gx, gy, m, p = Plane(w1, h1), Plane(w1, h1, "S16"), Plane(w1, h1, "S16"), Plane(w1, h1, "S16"), Plane(w1, h1)

    def chain():
        call("Gaussian_U8_U8_3x3", w1, h1, g, g.stride, s, s.stride)
        call("Sobel_S16S16_U8_3x3_GXY", w1, h1, gx, gx.stride, gy, gy.stride, g, g.stride)
        call("Magnitude_S16_S16S16", w1, h1, m, m.stride, gx, gx.stride, gy, gy.stride)
        call("Phase_U8_S16S16", w1, h1, p, p.stride, gx, gx.stride, gy, gy.stride)
    ms = timed(chain)
    out["chain1_1920x1080_per_node"] = bw("chain1", ms, w1 * h1, 8, w1 * h1 / (ms / 1e3) / 1e6, "Mpix/s")
    out["chain1_1920x1080_per_node"]["how"] = "four launches: Gaussian, Sobel GXY, Magnitude, Phase (blurred plane stored)"
    out["chain1_1920x1080"] = measure_chain1_graph(s, timed, bw)
    # warps and remap at 3840x2160
    src = synth.scene_u8(W, H, 1)
    sp, dp = Plane.from_array(src), Plane(W, H)
    aff, per = synth.warp_matrices(W, H, 1)
    aff_c, per_c = np.ascontiguousarray(aff), np.ascontiguousarray(per)
    ms = timed(lambda: call("WarpAffine_U8_U8_Bilinear_Constant", W, H, dp, dp.stride, W, H, sp, sp.stride,
                            aff_c.ctypes.data_as(ctypes.c_void_p), ctypes.c_uint8(0)))
    out["warp_affine_3840x2160"] = bw("warp_affine", ms, W * H, 2, W * H / (ms / 1e3) / 1e6, "Mpix/s")
    ms = timed(lambda: call("WarpPerspective_U8_U8_Bilinear_Constant", W, H, dp, dp.stride, W, H, sp, sp.stride,
                            per_c.ctypes.data_as(ctypes.c_void_p), ctypes.c_uint8(0)))
    out["warp_perspective_3840x2160"] = bw("warp_perspective", ms, W * H, 2, W * H / (ms / 1e3) / 1e6, "Mpix/s")
    fmap = torch.from_numpy(synth.barrel_map(W, H, 0.15)).cuda()
    fixed = torch.zeros((H, W * 4), dtype=torch.uint8, device="cuda")
    call("Remap_FixedMap_F32", W, H, ctypes.c_void_p(fixed.data_ptr()), W * 4, ctypes.c_void_p(fmap.data_ptr()), W * 8, W, H)
    ms = timed(lambda: call("Remap_U8_U8_Bilinear", W, H, dp, dp.stride, W, H, sp, sp.stride, sp.stride * H, ctypes.c_void_p(fixed.data_ptr()), W * 4))
    out["remap_3840x2160"] = bw("remap", ms, W * H, 6, W * H / (ms / 1e3) / 1e6, "Mpix/s")
    # the further kernel families at 4K
    ms = timed(lambda: call("WarpPerspective_U8_U8_Nearest_Constant", W, H, dp, dp.stride, W, H, sp, sp.stride,
                            per_c.ctypes.data_as(ctypes.c_void_p), ctypes.c_uint8(0)))
    out["warp_perspective_nearest_3840x2160"] = bw("warp_perspective_nearest", ms, W * H, 2, W * H / (ms / 1e3) / 1e6, "Mpix/s")
    u16 = Plane(W, H, "U16")
    ms = timed(lambda: call("CannySobel_U16_U8_5x5_L2NORM", W, H, u16, u16.stride, sp, sp.stride))
    out["canny_sobel_5x5_l2_3840x2160"] = bw("canny_sobel_5x5_l2", ms, W * H, 3, W * H / (ms / 1e3) / 1e6, "Mpix/s")
    # the whole Canny node (gradient 5, L2, hysteresis 80/100): gradient plane, suppression + double
    # threshold, edge trace; 1 B in + 1 B out per pixel
    edges, cstack = Plane(W, H), torch.zeros(4 * W * H, dtype=torch.uint8, device="cuda")

    def canny_node():
        call("CannySobel_U16_U8_5x5_L2NORM", W, H, u16, u16.stride, sp, sp.stride)
        call("CannySuppThreshold_U8XY_U16_3x3", W, H, edges, edges.stride, u16, u16.stride, ctypes.c_void_p(cstack.data_ptr()),
             W * H, 80, 100)
        call("CannyEdgeTrace_U8_U8XY", W, H, edges, edges.stride, ctypes.c_void_p(cstack.data_ptr()), W * H)
    ms = timed(canny_node)
    out["canny_5x5_l2_node_3840x2160"] = bw("canny_node", ms, W * H, 2, W * H / (ms / 1e3) / 1e6, "Mpix/s")
    out["canny_5x5_l2_node_3840x2160"]["how"] = "CannySobel 5x5 L2 + CannySuppThreshold + CannyEdgeTrace (tiled), one node"
    K = __import__("mivisionx_amd").kernels()
    K.mvx_FastCorners_workspace_size.restype = ctypes.c_size_t
    fws = int(K.mvx_FastCorners_workspace_size(W, H))
    fbuf = torch.zeros(fws + 20000 * 28 + 64, dtype=torch.uint8, device="cuda")
    fp = fbuf.data_ptr()
    ms = timed(lambda: call("FastCorners_XY_U8_Supression", 20000, ctypes.c_void_p(fp + fws), ctypes.c_void_p(fp + fws + 20000 * 28),
                            None, W, H, sp, sp.stride, 20.0, ctypes.c_void_p(fp), ctypes.c_size_t(fws)))
    out["fast_corners_nonmax_3840x2160"] = bw("fast_corners", ms, W * H, 2, W * H / (ms / 1e3) / 1e6, "Mpix/s")
    yp, up, vp_, rgbp = Plane(W, H), Plane(W // 2, H // 2), Plane(W // 2, H // 2), Plane(W, H, "RGB")
    ms = timed(lambda: call("ColorConvert_RGB_IYUV", W, H, rgbp, rgbp.stride, yp, yp.stride, up, up.stride, vp_, vp_.stride))
    out["color_rgb_from_iyuv_3840x2160"] = bw("rgb_from_iyuv", ms, W * H, 4.5, W * H / (ms / 1e3) / 1e6, "Mpix/s")
    rgbs = Plane(W, H, "RGB")
    ms = timed(lambda: call("Remap_U24_U24_Bilinear", W, H, rgbp, rgbp.stride, W, H, rgbs, rgbs.stride, ctypes.c_void_p(fixed.data_ptr()),
                            W * 4))
    out["remap_rgb_3840x2160"] = bw("remap_rgb", ms, W * H, 10, W * H / (ms / 1e3) / 1e6, "Mpix/s")
    # pyramidal LK at 3840x2160, 10,000 keypoints, on the two 4-level GaussianPyramid nodes of the graph
    # (level 0 a channel copy, each further level ScaleGaussianHalf 5x5 of the one before)
    old, new, _ = synth.lk_pair(W, H, 1)
    dims = pyramid_dims(W, H, 4)
    imgs = [Plane.from_array(a) for a in (old, new)]
    pyrs = [[Plane(w, h) for (w, h) in dims] for _ in range(2)]

    def pyramids():
        for img, lv in zip(imgs, pyrs):
            call("ChannelCopy_U8_U8", W, H, lv[0], lv[0].stride, img, img.stride)
            for k in range(1, len(dims)):
                call("ScaleGaussianHalf_U8_U8_5x5", dims[k][0], dims[k][1], lv[k], lv[k].stride, dims[k - 1][0], dims[k - 1][1],
                     lv[k - 1], lv[k - 1].stride)
    pyramids()
    tr = PyrLKTracker(pyrs[0], pyrs[1], 10000)
    tr.set_points(synth.lk_keypoints(W, H, 10000, 1))
    ms_lk = timed(lambda: tr.run(0xF002, 0.01, 20, False, 9))
    ms_pyr = timed(pyramids)

    def graph():
        pyramids()
        tr.run(0xF002, 0.01, 20, False, 9)
    ms = timed(graph)
    out["optical_flow_pyr_lk_4levels_10000pts"] = {
        "ms": round(ms, 4), "value": round(10000 / (ms / 1e3), 1), "unit": "keypoints/s",
        "includes": "both 4-level GaussianPyramid nodes (3840x2160 copy + 3 ScaleGaussianHalf 5x5 levels each) and the LK node",
        "lk_only_ms": round(ms_lk, 4), "pyramids_ms": round(ms_pyr, 4),
        "bound": "latency (one 64-lane workgroup per keypoint, sequential Newton steps)"}
    return out


def measure_chain1_graph(src, timed, bw, reps=200):
    """The 1920x1080 Gaussian (virtual) -> Sobel (gx, gy stored) -> Magnitude + Phase graph through the
    OpenVX engine: the chain supernode runs it as one launch.  Reported: the wall time of one
    vxProcessGraph (it synchronises), and that launch's device time -- the same mvx_ChainBatch kernel,
    plan and tile the engine chose, launched back to back and timed with HIP events on its stream."""
    import ctypes as c
    import torch
    from mivisionx_amd import openvx as vx
    from mivisionx_amd.device import Plane
    from mivisionx_amd.pipelines import CHAIN_TILES, ChainBatch
    w, h = src.width, src.height
    L = vx.lib()
    ctx = L.vxCreateContext()
    messages = []
    cb = vx.LOG_CB(lambda ctx_, ref, status, text: messages.append(text.decode()))
    L.vxRegisterLogCallback(ctx, cb, 0)
    addr = vx.Addressing(w, h, 1, src.stride, 1024, 1024, 1, 1, 0)
    img = L.vxCreateImageFromHandle(ctx, vx.DF["U8"], c.byref(addr), (c.c_void_p * 1)(src.t.data_ptr()), vx.VX_MEMORY_TYPE_HIP)
    g = L.vxCreateGraph(ctx)
    blur = L.vxCreateVirtualImage(g, 0, 0, vx.DF["VIRT"])
    o = [vx.create_image(ctx, w, h, f) for f in ("S16", "S16", "S16", "U8")]
    assert L.vxGaussian3x3Node(g, img, blur) and L.vxSobel3x3Node(g, blur, o[0], o[1])
    assert L.vxMagnitudeNode(g, o[0], o[1], o[2]) and L.vxPhaseNode(g, o[0], o[1], o[3])
    vx.check(L.vxVerifyGraph(g), "vxVerifyGraph")
    note = next((m.strip() for m in messages if "chain batch supernode" in m), "")
    for _ in range(5):
        vx.check(L.vxProcessGraph(g))
    t = time.perf_counter()
    for _ in range(reps):
        vx.check(L.vxProcessGraph(g))
    wall_ms = (time.perf_counter() - t) / reps * 1e3
    tile = next((i for i, (tw, th) in enumerate(CHAIN_TILES) if f"tile {tw}x{th}" in note), 4)
    cbatch = ChainBatch([src], with_half=False).prepare(tile)
    k_ms = timed(cbatch.run)
    rec = bw("chain1_graph", k_ms, w * h, 8, w * h / (k_ms / 1e3) / 1e6, "Mpix/s")
    rec.update({"how": "vxProcessGraph: the chain supernode, one launch (device time of that launch, HIP events)",
                "engine_log": note, "vxProcessGraph_wall_ms": round(wall_ms, 4),
                "vxProcessGraph_wall_value": round(w * h / (wall_ms / 1e3) / 1e6, 1)})
    gg, cc = c.c_void_p(g), c.c_void_p(ctx)
    L.vxReleaseGraph(c.byref(gg))
    L.vxReleaseContext(c.byref(cc))
    return rec


def measure_stitch(reps=10):
    """The stitch variant of the uneven-batch configuration: 6 RGB cameras of 1920x1080 (scene seeds
    4000+c) remapped through seeded F32 maps into a 7680x3840 equirectangular frame and merged with
    Loom's camera-selection maps and 64-px linear blend weights.  One pass (each output pixel samples
    only the cameras its group selects) against Loom's two passes (a U24 remap per camera into a
    stacked frame, then the merge), each timed with HIP events on its launch stream."""
    import torch
    from mivisionx_amd import synth
    from mivisionx_amd.pipelines import LoomStitch
    st = LoomStitch(synth.stitch_rig())
    torch.cuda.synchronize()

    def timed(fn):
        for _ in range(3):
            fn()
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
        for a, b in ev:
            a.record()
            fn()
            b.record()
        torch.cuda.synchronize()
        t = sorted(a.elapsed_time(b) for a, b in ev)
        return t[len(t) // 2]

    fused_ms = timed(st.run_fused)
    fused_out = st.out.t.clone()
    two_ms = timed(st.run_two_pass)
    same = bool(torch.equal(fused_out, st.out.t))
    alg = st.algorithmic_bytes()
    px = st.W * st.H
    gbs = alg / (fused_ms / 1e3) / 1e9
    return {
        "workload": "6 x 1920x1080 RGB cameras -> 7680x3840 RGB equirectangular, bilinear 13.3 fixed-point tables from F32 maps, "
                    "Loom camera-selection merge, 64-px linear overlaps",
        "one_pass": {"ms": round(fused_ms, 4), "value": round(px / (fused_ms / 1e3) / 1e6, 1), "unit": "output Mpix/s",
                     "kernel": "mvx_Stitch_Loom_RGB", "algorithmic_bytes": alg, "achieved_GBps": round(gbs, 1),
                     "frac": round(gbs / HBM_PEAK_GBS, 4), "bound": "hbm"},
        "two_pass": {"ms": round(two_ms, 4), "value": round(px / (two_ms / 1e3) / 1e6, 1), "unit": "output Mpix/s",
                     "how": "6 x mvx_Remap_U24_U24_Bilinear_Constant into a stacked frame + mvx_StitchMerge_Loom_RGB (Loom's passes)"},
        "one_pass_over_two_pass": round(two_ms / fused_ms, 2),
        "outputs_identical": same,
    }


def measure_batched_uneven(reps=5):
    """256 U8 images of uneven sizes, each Gaussian3x3 -> Sobel3x3 -> Magnitude + Phase plus ScaleImage
    to half size, through the OpenVX engine: one graph holding all 256 chains, run (a) with the
    chain-batch supernode -- one persistent launch replayed from a HIP graph -- and (b) with supernode
    merging off -- 1280 per-node launches, the reference executor's model.  Wall time per
    vxProcessGraph (it synchronises); the batch kernel also timed alone with HIP events."""
    import ctypes as c
    import numpy as np
    import torch
    from mivisionx_amd import openvx as vx, synth
    from mivisionx_amd.device import Plane
    from mivisionx_amd.pipelines import ChainBatch

    L = vx.lib()
    sizes = synth.batch_sizes()
    px = sum(w * h for w, h in sizes)
    planes = [Plane(w, h, "U8") for (w, h) in sizes]
    for i, (p, (w, h)) in enumerate(zip(planes, sizes)):
        p.t[:, :w].copy_(synth.scene_u8_device(w, h, seed=i))
    torch.cuda.synchronize()
    ctx = L.vxCreateContext()
    messages = []
    cb = vx.LOG_CB(lambda ctx_, ref, status, text: messages.append(text.decode()))
    L.vxRegisterLogCallback(ctx, cb, 0)
    srcs = []
    for p in planes:
        addr = vx.Addressing(p.width, p.height, 1, p.stride, 1024, 1024, 1, 1, 0)
        srcs.append(L.vxCreateImageFromHandle(ctx, vx.DF["U8"], c.byref(addr), (c.c_void_p * 1)(p.t.data_ptr()), vx.VX_MEMORY_TYPE_HIP))

    def build(merge):
        g = L.vxCreateGraph(ctx)
        if not merge:
            flags = c.c_uint32(vx.AGO_GRAPH_OPTIMIZER_FLAG_NO_SUPERNODE_MERGE)
            vx.check(L.vxSetGraphAttribute(g, vx.VX_GRAPH_ATTRIBUTE_AMD_OPTIMIZER_FLAGS, c.byref(flags), 4))
        outs = []
        for s, (w, h) in zip(srcs, sizes):
            blur = L.vxCreateVirtualImage(g, 0, 0, vx.DF["VIRT"])
            o = [vx.create_image(ctx, w, h, f) for f in ("S16", "S16", "S16", "U8")] + [vx.create_image(ctx, w // 2, h // 2, "U8")]
            assert L.vxGaussian3x3Node(g, s, blur) and L.vxSobel3x3Node(g, blur, o[0], o[1])
            assert L.vxMagnitudeNode(g, o[0], o[1], o[2]) and L.vxPhaseNode(g, o[0], o[1], o[3])
            assert L.vxScaleImageNode(g, s, o[4], vx.VX_INTERPOLATION_BILINEAR)
            outs.append(o)
        t = time.perf_counter()
        vx.check(L.vxVerifyGraph(g), "vxVerifyGraph")
        return g, outs, time.perf_counter() - t

    def wall(g):
        for _ in range(2):
            vx.check(L.vxProcessGraph(g))
        t = time.perf_counter()
        for _ in range(reps):
            vx.check(L.vxProcessGraph(g))
        return (time.perf_counter() - t) / reps * 1e3

    gb, ob, verify_b = build(True)
    note = next((m.strip() for m in messages if "chain batch supernode" in m), "")
    ms_b = wall(gb)
    gp, op, verify_p = build(False)
    ms_p = wall(gp)
    same = all(np.array_equal(vx.read_image(ob[i][k]), vx.read_image(op[i][k])) for i in (0, 255) for k in range(5))
    # the batch kernel alone (C ABI, the engine's tile choice), HIP events on its launch stream
    cbatch = ChainBatch(planes)
    tile = next(i for i, t in enumerate(((16, 16), (32, 16), (32, 32), (64, 16), (64, 32))) if f"tile {t[0]}x{t[1]}" in note) \
        if "tile" in note else 4
    cbatch.prepare(tile)
    cbatch.capture()
    for _ in range(2):
        cbatch.launch_graph()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        cbatch.launch_graph()
    b.record()
    torch.cuda.synchronize()
    k_ms = a.elapsed_time(b) / reps
    alg = int(px * 8.25)
    for g in (gb, gp):
        gg = c.c_void_p(g)
        L.vxReleaseGraph(c.byref(gg))
    cc = c.c_void_p(ctx)
    L.vxReleaseContext(c.byref(cc))
    return {
        "workload": "256 U8 images, W uniform in {320..4096}, H in {240..2160} (multiples of 16), seeded scene generator; "
                    "Gaussian3x3 -> Sobel3x3 -> Magnitude + Phase (blurred image virtual) + ScaleImage bilinear to half size",
        "images": len(sizes), "mpix": round(px / 1e6, 2),
        "batched": {"ms": round(ms_b, 3), "value": round(px / (ms_b / 1e3) / 1e6, 1), "unit": "Mpix/s",
                    "how": "vxProcessGraph, chain-batch supernode: one persistent launch replayed from a HIP graph",
                    "engine_log": note, "verify_s": round(verify_b, 2)},
        "per_image": {"ms": round(ms_p, 3), "value": round(px / (ms_p / 1e3) / 1e6, 1), "unit": "Mpix/s",
                      "how": "vxProcessGraph with supernode merge off: 5 launches per image (1280 per run)",
                      "verify_s": round(verify_p, 2)},
        "batched_over_per_image": round(ms_p / ms_b, 2),
        "outputs_identical": bool(same),
        "kernel": {"name": "k_chain_batch", "ms": round(k_ms, 4), "algorithmic_bytes": alg,
                   "achieved_GBps": round(alg / (k_ms / 1e3) / 1e9, 1), "frac": round(alg / (k_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
                   "bytes_per_px": "1 in + 4 gx/gy + 2 magnitude + 1 phase + 0.25 half = 8.25"},
    }


def measure_alignment(pairs=64, steps=200):
    """64 pairs of 1024x1024 F32 images: a coarse rotation/translation grid on the top pyramid level
    picks each pair's start, then Adam on the 3-level pyramid L2 loss through the fused warp + loss +
    adjoint kernel, one captured HIP graph per step; time per step and the recovered homography error"""
    import torch
    from mivisionx_amd import align
    src, tgt, Ht = align.make_pairs(pairs, 1024, 1024)
    al = align.Aligner(src, tgt)
    px_step = pairs * sum(a.shape[1] * a.shape[2] for a in al.ps)
    # fused kernel alone, HIP events on its stream (level 0)
    Hl = align.level_H(align.H_from_p(torch.zeros((pairs, 8), dtype=torch.float64, device="cuda"), 1024, 1024), 0)
    lv = torch.zeros(pairs, dtype=torch.float64, device="cuda")
    gH = torch.zeros((pairs, 9), dtype=torch.float64, device="cuda")
    for _ in range(2):
        al.l2[0](Hl, 1.0, lv, gH, False)
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(5):
        al.l2[0](Hl, 1.0, lv, gH, False)
    b.record()
    torch.cuda.synchronize()
    k_ms = a.elapsed_time(b) / 5
    t = time.perf_counter()
    p0 = al.coarse_init()
    torch.cuda.synchronize()
    init_ms = (time.perf_counter() - t) * 1e3
    p = p0.clone().requires_grad_(True)
    opt = torch.optim.Adam([p], lr=4e-3, capturable=True)
    graph = al.capture_step(p, opt)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(steps):
        graph.replay()
    torch.cuda.synchronize()
    step_ms = (time.perf_counter() - t) / steps * 1e3
    err = align.corner_error(align.H_from_p(p.detach(), 1024, 1024), Ht, 1024, 1024)
    # the same steps launched eagerly, for comparison
    q = p0.clone().requires_grad_(True)
    opt_e = torch.optim.Adam([q], lr=4e-3)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(20):
        _, g = al.loss_and_grad(q.detach())
        q.grad = g
        opt_e.step()
    torch.cuda.synchronize()
    eager_ms = (time.perf_counter() - t) / 20 * 1e3
    alg = pairs * 1024 * 1024 * 8
    return {
        "workload": f"{pairs} pairs of 1024x1024 F32, 3-level pyramid L2, coarse grid start, {steps} Adam steps",
        "ms_per_step": round(step_ms, 4), "value": round(px_step / (step_ms / 1e3) / 1e6, 1), "unit": "Mpix/s (pyramid pixels)",
        "how": "one captured HIP graph per Adam step (3 fused warp+L2+adjoint launches, their reductions, chain rule, optimiser)",
        "eager_ms_per_step": round(eager_ms, 4), "coarse_init_ms": round(init_ms, 1),
        "corner_error_px": {"mean": round(float(err.mean()), 4), "max": round(float(err.max()), 4),
                            "pairs_below_0.5px": int((err < 0.5).sum())},
        "kernel": {"name": "k_warp_f32_adjoint<L2> + k_reduce_partials, level 0", "ms": round(k_ms, 4), "algorithmic_bytes": alg,
                   "achieved_GBps": round(alg / (k_ms / 1e3) / 1e9, 1), "frac": round(alg / (k_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
                   "bytes_per_px": "4 source (bilinear gather) + 4 target"},
    }


def cpu_batched_uneven(images=3):
    """the reference CPU engine (one thread) on the first images of the uneven batch"""
    from mivisionx_amd import synth
    from pyoracle import Reference
    ref = Reference()
    tot_px, tot_s = 0, 0.0
    for i, (w, h) in enumerate(synth.batch_sizes()[:images]):
        _, sec = ref.chain_half(synth.scene_u8(w, h, seed=i), time_iters=2)
        tot_px += w * h
        tot_s += sec
    return {"value": round(tot_px / tot_s / 1e6, 2), "unit": "Mpix/s", "cores": 1, "kind": "reference",
            "sample": f"first {images} images of the uneven batch, Gaussian->Sobel->Magnitude+Phase + ScaleImage half, "
                      "one graph each, 2 timed runs"}


def cpu_nodes():
    """the reference CPU engine (oracle/_ref, one thread) on the node workloads, a few frames each"""
    from mivisionx_amd import synth
    from pyoracle import Reference
    ref = Reference()
    import numpy as np
    out = {}
    _, sec = ref.chain1(synth.uniform_u8(1920, 1080, 0), time_iters=3)
    out["chain1_1920x1080"] = {"ms": round(sec * 1e3, 2), "value": round(1920 * 1080 / sec / 1e6, 2), "unit": "Mpix/s"}
    src = synth.scene_u8(W, H, 1)
    aff, per = synth.warp_matrices(W, H, 1)
    _, sec = ref.warp(0, src, W, H, aff, True, 0, time_iters=2)
    out["warp_affine_3840x2160"] = {"ms": round(sec * 1e3, 2), "value": round(W * H / sec / 1e6, 2), "unit": "Mpix/s"}
    _, sec = ref.warp(1, src, W, H, per, True, 0, time_iters=2)
    out["warp_perspective_3840x2160"] = {"ms": round(sec * 1e3, 2), "value": round(W * H / sec / 1e6, 2), "unit": "Mpix/s"}
    _, sec = ref.remap(src, synth.barrel_map(W, H, 0.15), False, 0, time_iters=2)
    out["remap_3840x2160"] = {"ms": round(sec * 1e3, 2), "value": round(W * H / sec / 1e6, 2), "unit": "Mpix/s"}
    old, new, _ = synth.lk_pair(W, H, 1)
    kp = synth.lk_keypoints(W, H, 10000, 1)
    _, sec = ref.lk(old, new, 4, kp, kp, 0xF002, 0.01, 20, 0, 9, reps=2, timed=True)
    out["optical_flow_pyr_lk_4levels_10000pts"] = {"ms": round(sec * 1e3, 2), "value": round(10000 / sec, 1), "unit": "keypoints/s",
                                                   "note": "includes the two 4-level GaussianPyramid nodes of the graph"}
    return out


def cpu_baseline(rgb, budget_s=10.0):
    from pyoracle import Oracle, Reference
    try:
        ref = Reference()
        r = ref.graph2(rgb, DW, DH, LEVELS, PARAMS["lower"], PARAMS["upper"], PARAMS["k"], PARAMS["thr"],
                       PARAMS["dist"], PARAMS["capacity"], time_iters=1)
        iters = max(2, int(budget_s / max(r["sec_per_frame"], 1e-3)))
        r = ref.graph2(rgb, DW, DH, LEVELS, PARAMS["lower"], PARAMS["upper"], PARAMS["k"], PARAMS["thr"],
                       PARAMS["dist"], PARAMS["capacity"], time_iters=iters)
        spf, kind = r["sec_per_frame"], "reference"
        sample = f"{iters} frames of the 3840x2160 graph through the reference CPU engine (AGO, ago_haf_cpu*, -O3 -msse4.2)"
    except FileNotFoundError:
        orc = Oracle()
        t0 = time.perf_counter()
        iters = 0
        while time.perf_counter() - t0 < budget_s or iters < 2:
            orc.graph2(rgb, DW, DH, LEVELS, PARAMS["lower"], PARAMS["upper"], PARAMS["k"], PARAMS["thr"],
                       PARAMS["dist"], PARAMS["capacity"])
            iters += 1
        spf, kind = (time.perf_counter() - t0) / iters, "port"
        sample = f"{iters} frames of the 3840x2160 graph through the oracle's scalar C port"
    return {"value": round(W * H / spf / 1e6, 3), "unit": "Mpix/s", "cores": 1, "kind": kind, "sample": sample,
            "ms_per_frame": round(spf * 1e3, 2), "host": host_cpu(), "all_cores": cpu_port_threads(rgb)}


def host_cpu():
    """the host's CPU model and the cores this process may use"""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), model)
    except OSError:
        pass
    return {"cpu_model": model, "nproc": len(os.sched_getaffinity(0)), "os_cpu_count": os.cpu_count()}


def cpu_port_threads(rgb, max_threads=16):
    """the oracle's C port of the graph on every usable host core: one frame per thread, threads
    running concurrently (the port's C calls release the interpreter lock); frames per second of
    the whole pool.  The reference engine itself is single-threaded, so this is the port, not it."""
    from concurrent.futures import ThreadPoolExecutor
    from pyoracle import Oracle
    n = max(1, min(max_threads, len(os.sched_getaffinity(0))))
    orc = Oracle()

    def one(_):
        orc.graph2(rgb, DW, DH, LEVELS, PARAMS["lower"], PARAMS["upper"], PARAMS["k"], PARAMS["thr"], PARAMS["dist"],
                   PARAMS["capacity"])
    with ThreadPoolExecutor(n) as pool:
        list(pool.map(one, range(n)))          # first touch of every thread
        t = time.perf_counter()
        list(pool.map(one, range(2 * n)))
        sec = time.perf_counter() - t
    return {"value": round(2 * n * W * H / sec / 1e6, 3), "unit": "Mpix/s", "cores": n, "kind": "port",
            "sample": f"{2 * n} frames of the 3840x2160 graph through the oracle's C port, {n} threads, one frame each at a time"}


def traffic_from_profiles(stage, frames):
    """HBM bytes per launch of the dominant stage, from the committed rocprofv3 --pmc summary
    (profiles/k_front_4k_traffic.md: FETCH_SIZE and WRITE_SIZE per dispatch, scaled by the copy
    kernel's reading), per frame times the frames of one launch; None when no summary covers it"""
    import re
    path = os.path.join(ROOT, "profiles", "k_front_4k_traffic.md")
    if stage != "fused_tile_pass" or not os.path.exists(path):
        return None
    m = re.search(r"traffic_bytes_per_frame:\s*(\d+)", open(path).read())
    return int(int(m.group(1)) * frames) if m else None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--frames", type=int, default=16, help="4K frames per step (resident, distinct buffers)")
    ap.add_argument("--chunk", type=int, default=0,
                    help="frames per tile-pass launch (0: the library's default, all frames in one launch)")
    ap.add_argument("--cpu-budget", type=float, default=10.0, help="seconds of CPU baseline work")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-per-image", action="store_true")
    ap.add_argument("--no-nodes", action="store_true", help="skip the per-node measurements (1080p chain, 4K warps, remap, LK, FAST, Canny)")
    ap.add_argument("--no-batch", action="store_true", help="skip the uneven-size batch measurement")
    ap.add_argument("--no-align", action="store_true", help="skip the F32 alignment measurement")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run 3 launches of a plain 128 MiB device copy (ChannelCopy 16384x8192), the known-size "
                         "kernel that scales rocprofv3 FETCH_SIZE / WRITE_SIZE readings into bytes")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the vision graph runs on a single MI355X here (asked for --gpus {args.gpus}); "
              "DESIGN.md describes how the path would shard across GPUs.", file=sys.stderr)
        sys.exit(2)

    import torch
    from mivisionx_amd import kernels, synth
    from mivisionx_amd.pipelines import FusedVisionGraph, VisionGraph4K

    torch.cuda.set_device(0)
    K = kernels()
    if args.copy_calibration:
        from mivisionx_amd.device import Plane, call
        a, b = Plane(16384, 8192), Plane(16384, 8192)
        a.t.random_(0, 256)
        for _ in range(3):
            call("ChannelCopy_U8_U8", 16384, 8192, b, b.stride, a, a.stride)
        torch.cuda.synchronize()
        del a, b
    base = [torch.from_numpy(synth.scene_rgb(W, H, s)) for s in (0, 1)]
    G = FusedVisionGraph(W, H, DW, DH, LEVELS, frames=args.frames, **PARAMS)
    for f, fr in enumerate(G.frames):
        t = base[f % 2].cuda()
        if f >= 2:
            t = torch.roll(t, shifts=(37 * f, 53 * f), dims=(0, 1))
        fr.rgb.t[:, : 3 * W].copy_(t.reshape(H, 3 * W))
    torch.cuda.synchronize()

    def stream():
        return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)

    def run_stage(st, f0=0, n=None):
        n = len(G.frames) if n is None else n
        rc = K.mvx_VisionGraph_run_stage_range(stream(), ctypes.byref(G.params), f0, n, ctypes.c_void_p(G.ws.data_ptr()),
                                               ctypes.c_size_t(G.ws_size), st)
        if rc:
            raise RuntimeError(f"stage {st} returned {rc}")

    # warm-up with every stage bracketed by events (all frames per launch): find the dominant stage
    ev = {st: [] for st in range(len(STAGES))}
    for _ in range(max(1, args.warmup)):
        for st in range(len(STAGES)):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record(); run_stage(st); b.record()
            ev[st].append((a, b))
    torch.cuda.synchronize()
    stage_ms = {st: sum(a.elapsed_time(b) for a, b in v) / len(v) for st, v in ev.items()}
    dominant = max(stage_ms, key=stage_ms.get)

    # timed region: K steps of all frames in mvx_VisionGraph_run's schedule (default-priority streams);
    # every launch of the dominant stage bracketed by events on its stream
    nf = len(G.frames)
    chunk = min(args.chunk, nf) if args.chunk > 0 else nf
    dom_ev = []
    main_s, hyst_s, side_s = torch.cuda.current_stream(), torch.cuda.Stream(), torch.cuda.Stream()

    def timed_stage(st, f0, n):
        if st == dominant:
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record(); run_stage(st, f0, n); b.record()
            dom_ev.append((a, b, n))
        else:
            run_stage(st, f0, n)

    def pipelined_step():
        for f0 in range(0, nf, chunk):
            n = min(chunk, nf - f0)
            timed_stage(0, f0, n)
            hyst_s.wait_stream(main_s)
            side_s.wait_stream(main_s)
            with torch.cuda.stream(side_s):
                timed_stage(3, f0, n)
                timed_stage(4, f0, n)
            # the last chunk's hysteresis stays on the main stream, as in the library
            with torch.cuda.stream(hyst_s if f0 + n < nf else main_s):
                timed_stage(1, f0, n)
                timed_stage(2, f0, n)
        main_s.wait_stream(hyst_s)
        main_s.wait_stream(side_s)

    # untimed: the pipelined step itself, so the side streams' first use is not in the timed region
    for _ in range(max(1, args.warmup)):
        pipelined_step()
    dom_ev.clear()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record()
    for _ in range(args.steps):
        pipelined_step()
    e1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = e0.elapsed_time(e1)
    dom_ms = sum(a.elapsed_time(b) for a, b, _ in dom_ev) / len(dom_ev)
    dom_frames = sum(n for _, _, n in dom_ev) / len(dom_ev)
    frames = args.steps * nf
    value = frames * W * H / (gpu_ms / 1e3) / 1e6
    dom_bytes = stage_algorithmic_bytes(dominant) * dom_frames
    achieved = dom_bytes / (dom_ms / 1e3) / 1e9
    graph_bytes = graph_algorithmic_bytes()
    graph_gbs = graph_bytes * frames / (gpu_ms / 1e3) / 1e9
    corners0 = G.frames[0].results()["count"]

    out = {
        "metric": "graph Mpix/s at 4K U8 (achieved HBM GB/s %peak); batched vs per-image",
        "value": round(value, 2), "unit": "Mpix/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(gpu_ms / args.steps, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u8", "data": "synthetic",
        "config": {"workload": "3840x2160 RGB vision graph: ColorConvert->Y; ScaleImage 1920x1080 bilinear; "
                               "GaussianPyramid 5 HALF; Canny 3 L1 80/100; Harris 0.04/0.0005/5/3/3",
                   "frames_per_step": nf, "frames_per_tile_pass_launch": chunk, "width": W, "height": H,
                   "parallelism": "single-gpu"},
        "roofline": {"bound": "hbm", "kernel": STAGES[dominant], "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "algorithmic_bytes_per_launch": int(dom_bytes), "mean_launch_ms": round(dom_ms, 5),
                     "frames_per_launch": dom_frames, "traffic": traffic_from_profiles(STAGES[dominant], dom_frames)},
        "graph_hbm": {"algorithmic_bytes_per_frame": graph_bytes, "achieved": round(graph_gbs, 1), "unit": "GB/s",
                      "frac": round(graph_gbs / HBM_PEAK_GBS, 4)},
        "stage_ms_per_step": {STAGES[k]: round(v, 4) for k, v in stage_ms.items()},
        "corners_frame0": corners0, "wall_s": round(wall, 3),
    }

    if not args.no_per_image:
        # the reference's execution model on the same frames: one launch per node, frame after frame
        pg = VisionGraph4K(W, H, DW, DH, LEVELS, **PARAMS)
        steps_pi = max(2, args.steps // 4)
        for fr in G.frames[:2]:
            pg.run(fr.rgb)
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(steps_pi):
            for fr in G.frames:
                pg.run(fr.rgb)
        b.record()
        torch.cuda.synchronize()
        pi_ms = a.elapsed_time(b)
        pi_value = steps_pi * len(G.frames) * W * H / (pi_ms / 1e3) / 1e6
        out["per_image"] = {"value": round(pi_value, 2), "unit": "Mpix/s",
                            "what": "same frames, one launch per node and frame (no fusion, no batching)",
                            "batched_over_per_image": round(value / pi_value, 2)}
    if not args.no_nodes:
        out["nodes"] = measure_nodes()
    if not args.no_batch:
        out["batched_uneven"] = measure_batched_uneven()
    if not args.no_batch:
        out["stitch"] = measure_stitch()
    if not args.no_align:
        out["alignment"] = measure_alignment()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(base[0].numpy(), args.cpu_budget)
        if not args.no_nodes:
            try:
                out["cpu_baseline"]["nodes"] = cpu_nodes()
            except FileNotFoundError:
                pass
        if not args.no_batch:
            try:
                out["cpu_baseline"]["batched_uneven"] = cpu_batched_uneven()
            except FileNotFoundError:
                pass
    print(json.dumps(out))


if __name__ == "__main__":
    main()
