"""The stitch path on the GPU: per-camera RGB remap (three U8 remaps) then the weighted merge, against the
oracle's per-channel remap and merge, bit-exact."""
import ctypes

import numpy as np
import pytest
import torch

from mivisionx_amd import synth
from mivisionx_amd.device import Plane, call
from util import assert_same

pytestmark = pytest.mark.gpu


def test_two_camera_stitch(gpu, orc):
    sw, sh, w, h = 640, 480, 800, 400
    cams = [np.stack([synth.scene_u8(sw, sh, 10 * k + c) for c in range(3)], -1) for k in range(3)]
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float32)
    maps, warped, weights = [], [], []
    for k in range(3):
        mp = np.stack([(xx - 150.0 * k) * 0.9 + 20.0, yy * 1.1 + 3.0 * k], -1).astype(np.float32)
        maps.append(mp)
        wt = np.clip(255 - np.abs(xx - (200 + 200 * k)) * 1.2, 0, 255).astype(np.uint8)
        weights.append(wt)
    dev = []
    for k in range(3):
        s = Plane.from_array(cams[k], "RGB")
        fmap = torch.from_numpy(np.ascontiguousarray(maps[k])).cuda()
        fixed = torch.zeros((h, w * 4), dtype=torch.uint8, device="cuda")
        call("Remap_FixedMap_F32", w, h, ctypes.c_void_p(fixed.data_ptr()), w * 4, ctypes.c_void_p(fmap.data_ptr()), w * 8, sw, sh)
        d = Plane(w, h, "RGB")
        call("Remap_U24_U24_Bilinear_Constant", w, h, d, d.stride, sw, sh, s, s.stride, ctypes.c_void_p(fixed.data_ptr()), w * 4,
             ctypes.c_uint8(0))
        want = np.stack([orc.remap(np.ascontiguousarray(cams[k][..., c]), maps[k], True, 0) for c in range(3)], -1)
        assert_same(d.download(), want, f"camera {k} remap")
        warped.append(want)
        dev.append((s, d, fixed, Plane.from_array(weights[k])))
    out = Plane(w, h, "RGB")
    P = ctypes.c_void_p * 3
    U = ctypes.c_uint32 * 3
    call("StitchMerge_RGB", w, h, out, out.stride, 3, P(*[x[1].t.data_ptr() for x in dev]), U(*[x[1].stride for x in dev]),
         P(*[x[3].t.data_ptr() for x in dev]), U(*[x[3].stride for x in dev]))
    assert_same(out.download(), orc.stitch_merge(warped, weights), "stitch merge")


def test_loom_merge_random_selection(gpu, orc):
    """mvx_StitchMerge_Loom_RGB against the oracle on random selection maps: every code range, ids of 31
    and ids beyond the cameras, odd group counts"""
    from mivisionx_amd.device import Plane
    w, h, n = 200, 37, 5
    rng = np.random.default_rng(3)
    stacked = rng.integers(0, 256, (n * h, w, 3), dtype=np.uint8)
    wts = rng.integers(0, 256, (n * h, w), dtype=np.uint8)
    g = (w + 7) // 8
    id0 = rng.choice(np.array([0, 1, 2, 3, 4, 7, 31, 40, 128, 129, 130, 131, 132, 250], np.uint8), (h, g))
    id1 = rng.integers(0, 1 << 15, (h, g), dtype=np.uint16) & np.uint16(0x7FFF)
    id2 = rng.integers(0, 1 << 15, (h, g), dtype=np.uint16)
    ids = lambda a: np.where(((a[..., None] >> np.array([0, 5, 10])) & 31) >= n + 3, 31, (a[..., None] >> np.array([0, 5, 10])) & 31)
    id1 = (ids(id1) << np.array([0, 5, 10])).sum(-1).astype(np.uint16)
    id2 = (ids(id2) << np.array([0, 5, 10])).sum(-1).astype(np.uint16)
    out = Plane(w, h, "RGB")
    out.t.fill_(77)
    S, Wt = torch.from_numpy(stacked).cuda(), torch.from_numpy(wts).cuda()
    I0, I1, I2 = (torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (id0, id1.view(np.int16), id2.view(np.int16)))
    call("StitchMerge_Loom_RGB", w, h, out, out.stride, n, ctypes.c_void_p(S.data_ptr()), 3 * w, ctypes.c_void_p(Wt.data_ptr()), w,
         ctypes.c_void_p(I0.data_ptr()), g, ctypes.c_void_p(I1.data_ptr()), 2 * g, ctypes.c_void_p(I2.data_ptr()), 2 * g)
    want = orc.stitch_loom(stacked, wts, id0, id1, id2, dst=np.full((h, w, 3), 77, np.uint8))
    assert_same(out.download(), want, "Loom merge, random selection")


def test_loom_stitch_config_size(gpu, orc):
    """6 cameras of 1920x1080 RGB into a 7680x3840 equirectangular frame, linear weights over 64-pixel
    overlaps, Loom's selection maps: the one-pass stitch equals Loom's two passes on the GPU and the
    oracle (remap through the same 13.3 tables, then the merge)"""
    from mivisionx_amd.pipelines import LoomStitch
    print("stitch rig 6 x 1920x1080 -> 7680x3840", flush=True)
    rig = synth.stitch_rig()
    S = LoomStitch(rig)
    S.out.t.fill_(0)
    S.run_fused()
    one = S.results()
    S.out.t.fill_(0)
    S.run_two_pass()
    two = S.results()
    assert_same(one, two, "one-pass vs two-pass stitch")
    W, H = rig["W"], rig["H"]
    fixed = S.fixed.cpu().numpy().view(np.uint16).reshape(S.count * H, W, 2)
    args = (rig["weights"].cpu().numpy(), rig["camId0"].cpu().numpy(), rig["camId1"].cpu().numpy().view(np.uint16),
            rig["camId2"].cpu().numpy().view(np.uint16))
    want = orc.stitch_loom_remap(rig["images"], fixed, *args, workers=16)
    assert_same(one, want, "stitch vs oracle")
    sel = rig["camId0"].cpu().numpy()
    assert (sel < 31).any() and (sel == 31).any() and (sel > 31).any()
