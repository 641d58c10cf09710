"""GPU parity of WarpAffine, WarpPerspective and Remap (bilinear and nearest neighbour, UNDEFINED and
CONSTANT borders) and of the GPU-built fixed-point remap table: HIP kernels through the C ABI (and
through the vx API) versus the oracle, bit-exact including the padding pixels the 4-pixel stores
write, plus the reference engine's golden hashes."""
import ctypes
import json
import os

import numpy as np
import pytest
import torch

from mivisionx_amd import synth
from mivisionx_amd.device import Plane, call
from util import assert_same, sha

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SIZES = [(333, 77), (640, 480), (64, 48), (37, 29), (16, 5), (3840, 2160)]


def _fptr(m):
    m = np.ascontiguousarray(np.asarray(m, np.float32))
    return m, m.ctypes.data_as(ctypes.c_void_p)


def _padded(plane):
    """the whole rows, padding included"""
    return plane.t.cpu().numpy()


def _run_warp(kind, src_plane, dw, dh, m, border):
    d = Plane(dw, dh)
    keep, mp = _fptr(m)
    name = ("WarpAffine" if kind == 0 else "WarpPerspective") + "_U8_U8_Bilinear"
    if border is None:
        call(name, dw, dh, d, d.stride, src_plane.width, src_plane.height, src_plane, src_plane.stride, mp)
    else:
        call(name + "_Constant", dw, dh, d, d.stride, src_plane.width, src_plane.height, src_plane, src_plane.stride, mp,
             ctypes.c_uint8(border))
    return d


@pytest.mark.parametrize("w,h", SIZES)
def test_warps_match_oracle(gpu, orc, w, h):
    src = synth.scene_u8(w, h, 40)
    s = Plane.from_array(src)
    seeds = (1, 2) if w * h < 4_000_000 else (1,)
    for seed in seeds:
        aff, per = synth.warp_matrices(w, h, seed)
        for (dw, dh) in ((w, h), (w + 5, max(3, h - 3))):
            for border in (None, 0, 77):
                d = _run_warp(0, s, dw, dh, aff, border)
                want, _ = orc.warp_affine(src, dw, dh, aff, border is not None, border or 0, padded=True)
                assert_same(_padded(d), want, f"affine {w}x{h}->{dw}x{dh} seed {seed} border {border}")
                d = _run_warp(1, s, dw, dh, per, border)
                want, _ = orc.warp_perspective(src, dw, dh, per, border or 0, padded=True)
                assert_same(_padded(d), want, f"perspective {w}x{h}->{dw}x{dh} seed {seed} border {border}")


def test_warp_edge_matrices(gpu, orc):
    """the plain (no border test) loop, exact negative integer coordinates, huge coordinates"""
    w, h = 120, 64
    src = synth.uniform_u8(w, h, 41)
    s = Plane.from_array(src)
    mats = [np.array([0.5, 0.01, -0.01, 0.5, 20.3, 30.7], np.float32),   # maps inside the source
            np.array([1, 0, 0, 1, -1, -2], np.float32), np.array([1, 0, 0, 1, 0.5, 0.75], np.float32),
            np.array([1e9, 0, 0, 1, 0, 0], np.float32), np.array([1, 0, 0, 1, -3e9, 0], np.float32)]
    for m in mats:
        for border in (None, 3):
            d = _run_warp(0, s, w // 2, h // 2, m, border)
            want, _ = orc.warp_affine(src, w // 2, h // 2, m, border is not None, border or 0, padded=True)
            assert_same(_padded(d), want, f"affine {m.tolist()} {border}")
            p = np.array([m[0], m[1], 0, m[2], m[3], 0, m[4], m[5], 1], np.float32)
            d = _run_warp(1, s, w // 2, h // 2, p, border)
            want, _ = orc.warp_perspective(src, w // 2, h // 2, p, border or 0, padded=True)
            assert_same(_padded(d), want, f"perspective {p.tolist()} {border}")
    # a perspective with a zero denominator column (z = inf)
    p = np.array([1, 0, -0.01, 0, 1, 0, 0, 0, 0.3], np.float32)
    d = _run_warp(1, s, w, h, p, 11)
    want, _ = orc.warp_perspective(src, w, h, p, 11, padded=True)
    assert_same(_padded(d), want, "perspective with z = 0 column")


def test_warps_staging_limits(gpu, orc):
    """transforms whose tile footprints exceed the staged box (zoom-out by 3 and by 40, a 45 degree
    rotation, a perspective whose horizon crosses the output) and a source the staged kernels cannot
    read as dwords (a base one byte past a 16-byte boundary, stride not a multiple of 4)"""
    w, h = 700, 300
    src = synth.scene_u8(w, h, 43)
    s = Plane.from_array(src)
    c, sn = np.cos(np.pi / 4), np.sin(np.pi / 4)
    mats = [np.array([3, 0, 0, 3, 1.5, 0.5], np.float32), np.array([40, 0, 0, 40, -3, -2], np.float32),
            np.array([c, sn, -sn, c, 200.25, 10.5], np.float32)]
    for m in mats:
        for border in (None, 5):
            d = _run_warp(0, s, w, h, m, border)
            want, _ = orc.warp_affine(src, w, h, m, border is not None, border or 0, padded=True)
            assert_same(_padded(d), want, f"affine {m.tolist()} {border}")
    p = np.array([1, 0.02, 0.002, 0.1, 1, 0.0005, 5, 3, -0.4], np.float32)
    for border in (None, 9):
        d = _run_warp(1, s, w, h, p, border)
        want, _ = orc.warp_perspective(src, w, h, p, border or 0, padded=True)
        assert_same(_padded(d), want, f"perspective across the horizon {border}")
    # unaligned source: a packed copy at byte offset 1 with stride w + 1 (odd)
    st = w + 1
    raw = torch.zeros(h * st + 64, dtype=torch.uint8, device="cuda")
    raw[1:1 + h * st].view(h, st)[:, :w].copy_(torch.from_numpy(src).cuda())
    aff, per = synth.warp_matrices(w, h, 3)
    padded = np.zeros((h, st), np.uint8)
    padded[:, :w] = src
    for kind, m in ((0, aff), (1, per)):
        d = Plane(w, h)
        keep, mp = _fptr(m)
        name = ("WarpAffine" if kind == 0 else "WarpPerspective") + "_U8_U8_Bilinear_Constant"
        call(name, w, h, d, d.stride, w, h, ctypes.c_void_p(raw.data_ptr() + 1), st, mp, ctypes.c_uint8(7))
        fn = orc.warp_affine if kind == 0 else orc.warp_perspective
        want, _ = (fn(src, w, h, m, True, 7, padded=True, src_stride=st) if kind == 0 else fn(src, w, h, m, 7, padded=True, src_stride=st))
        assert_same(_padded(d), want, f"unaligned source kind {kind}")


def _remap(src_plane, mp, border, nearest=False):
    dh, dw = mp.shape[:2]
    fmap = torch.from_numpy(np.ascontiguousarray(mp)).cuda()
    fixed = torch.zeros((dh, dw * 4), dtype=torch.uint8, device="cuda")
    call("Remap_FixedMap_F32", dw, dh, ctypes.c_void_p(fixed.data_ptr()), dw * 4, ctypes.c_void_p(fmap.data_ptr()), dw * 8,
         src_plane.width, src_plane.height)
    d = Plane(dw, dh)
    name = "Remap_U8_U8_Nearest" if nearest else "Remap_U8_U8_Bilinear"
    size = src_plane.stride * src_plane.height
    if border is None:
        call(name, dw, dh, d, d.stride, src_plane.width, src_plane.height, src_plane, src_plane.stride, size,
             ctypes.c_void_p(fixed.data_ptr()), dw * 4)
    elif nearest:
        call(name + "_Constant", dw, dh, d, d.stride, src_plane.width, src_plane.height, src_plane, src_plane.stride, size,
             ctypes.c_void_p(fixed.data_ptr()), dw * 4, ctypes.c_uint8(border))
    else:  # the reference's Bilinear_Constant form takes no buffer size
        call(name + "_Constant", dw, dh, d, d.stride, src_plane.width, src_plane.height, src_plane, src_plane.stride,
             ctypes.c_void_p(fixed.data_ptr()), dw * 4, ctypes.c_uint8(border))
    return d, fixed.cpu().numpy().view(np.uint16).reshape(dh, dw, 2)


@pytest.mark.parametrize("w,h", SIZES)
def test_remap_matches_oracle(gpu, orc, w, h):
    src = synth.scene_u8(w, h, 42)
    s = Plane.from_array(src)
    for k1 in ((-0.2, 0.15) if w * h < 4_000_000 else (-0.2,)):
        mp = synth.barrel_map(w, h, k1)
        if w * h < 4_000_000:
            mp[::5, ::3] = [-2.0, 1.0]       # points outside the source
        for border in (None, 200):
            d, fixed = _remap(s, mp, border)
            assert_same(fixed, orc.remap_fixed(mp, w, h), f"fixed map {w}x{h} k1 {k1}")
            assert_same(_padded(d), orc.remap(src, mp, border is not None, border or 0, padded=True),
                        f"remap {w}x{h} k1 {k1} border {border}")


def _run_warp_nn(kind, src_plane, dw, dh, m, border, valid=None):
    d = Plane(dw, dh)
    keep, mp = _fptr(m)
    name = ("WarpAffine" if kind == 0 else "WarpPerspective") + "_U8_U8_Nearest"
    args = (dw, dh, d, d.stride, src_plane.width, src_plane.height, src_plane, src_plane.stride, mp)
    if border is None:
        call(name, *args)
    elif kind == 0:
        rect = (ctypes.c_uint32 * 4)(*valid) if valid is not None else None
        call(name + "_Constant", *args, ctypes.c_uint8(border), rect)
    else:
        call(name + "_Constant", *args, ctypes.c_uint8(border))
    return d


@pytest.mark.parametrize("w,h", SIZES)
def test_nearest_warps_match_oracle(gpu, orc, w, h):
    src = synth.scene_u8(w, h, 44)
    s = Plane.from_array(src)
    for seed in ((1, 2) if w * h < 4_000_000 else (1,)):
        aff, per = synth.warp_matrices(w, h, seed)
        for (dw, dh) in ((w, h), (w + 5, max(3, h - 3)), (w + 4, h + 1)):
            for border in (None, 0, 77):
                d = _run_warp_nn(0, s, dw, dh, aff, border)
                assert_same(_padded(d), orc.warp_affine_nearest(src, dw, dh, aff, border, padded=True),
                            f"affine nearest {w}x{h}->{dw}x{dh} seed {seed} border {border}")
                d = _run_warp_nn(1, s, dw, dh, per, border)
                assert_same(_padded(d), orc.warp_perspective_nearest(src, dw, dh, per, border, padded=True),
                            f"perspective nearest {w}x{h}->{dw}x{dh} seed {seed} border {border}")
    # a valid rectangle narrower than the source masks the CONSTANT form further
    aff = np.array([0.9, 0.05, -0.04, 0.92, 3.3, 2.7], np.float32)
    valid = (2, 3, max(3, w - 5), max(4, h - 4))
    d = _run_warp_nn(0, s, w, h, aff, 9, valid)
    assert_same(_padded(d), orc.warp_affine_nearest(src, w, h, aff, 9, valid=valid, padded=True), f"affine nearest valid rect {w}x{h}")


@pytest.mark.parametrize("w,h", SIZES)
def test_remap_nearest_matches_oracle(gpu, orc, w, h):
    src = synth.scene_u8(w, h, 45)
    s = Plane.from_array(src)
    for k1 in ((-0.2, 0.15) if w * h < 4_000_000 else (-0.2,)):
        mp = synth.barrel_map(w, h, k1)
        if w * h < 4_000_000:
            mp[::5, ::3] = [-2.0, 1.0]
        for border in (None, 3, 200):
            d, _ = _remap(s, mp, border, nearest=True)
            want, unp = orc.remap_nearest(src, mp, border is not None, border or 0, padded=True)
            got = _padded(d)
            got[:, :w][unp] = 0
            want[:, :w][unp] = 0
            assert_same(got, want, f"remap nearest {w}x{h} k1 {k1} border {border}")


def test_warp_and_remap_golden(gpu, orc):
    with open(os.path.join(ROOT, "tests", "golden", "vision_golden_warp.json")) as f:
        g = json.load(f)
    n = 0
    for c in g["cases"]:
        if c["w"] != 1920:
            continue
        src = synth.scene_u8(c["w"], c["h"], g["scene_seed"])
        s = Plane.from_array(src)
        if "affine" in c:
            aff = np.array(c["affine"], np.float32)
            d = _run_warp(0, s, c["dw"], c["dh"], aff, c["border"])
            _, unp = orc.warp_affine(src, c["dw"], c["dh"], aff, c["border"] is not None, c["border"] or 0)
            got = d.download()
            got[unp] = 0
            assert sha(got) == c["affine_sha256"], ("affine", c["seed"], c["border"])
            n += 1
        else:
            d, _ = _remap(s, synth.barrel_map(c["w"], c["h"], c["k1"]), c["border"])
            assert sha(d.download()) == c["remap_sha256"], ("remap", c["k1"], c["border"])
            n += 1
    assert n >= 8


def test_warp_remap_through_vx_api(gpu, orc):
    from mivisionx_amd import openvx as vx
    L = vx.lib()
    ctx = L.vxCreateContext()
    try:
        w, h = 300, 200
        src = synth.scene_u8(w, h, 43)
        a = vx.create_image(ctx, w, h, "U8")
        vx.write_image(a, src)
        aff, per = synth.warp_matrices(w, h, 5)
        out = vx.create_image(ctx, w, h, "U8")
        for kind, m in ((0, aff), (1, per)):
            mat = L.vxCreateMatrix(ctx, vx.VX_TYPE_FLOAT32, 2 if kind == 0 else 3, 3)
            keep = np.ascontiguousarray(m)
            assert L.vxCopyMatrix(mat, keep.ctypes.data_as(ctypes.c_void_p), vx.VX_WRITE_ONLY, vx.VX_MEMORY_TYPE_HOST) == 0
            for border, interp in ((None, vx.VX_INTERPOLATION_BILINEAR), (90, vx.VX_INTERPOLATION_BILINEAR),
                                   (None, vx.VX_INTERPOLATION_NEAREST_NEIGHBOR), (90, vx.VX_INTERPOLATION_NEAREST_NEIGHBOR)):
                g = L.vxCreateGraph(ctx)
                node = (L.vxWarpAffineNode if kind == 0 else L.vxWarpPerspectiveNode)(g, a, mat, interp, out)
                if border is not None:
                    b = vx.Border(vx.VX_BORDER_CONSTANT)
                    b.constant_value.U8 = border
                    assert L.vxSetNodeAttribute(node, vx.VX_NODE_BORDER, ctypes.byref(b), ctypes.sizeof(b)) == 0
                assert L.vxProcessGraph(g) == 0
                if interp == vx.VX_INTERPOLATION_NEAREST_NEIGHBOR:
                    want = (orc.warp_affine_nearest if kind == 0 else orc.warp_perspective_nearest)(src, w, h, m, border)
                elif kind == 0:
                    want, _ = orc.warp_affine(src, w, h, m, border is not None, border or 0)
                else:
                    want, _ = orc.warp_perspective(src, w, h, m, border or 0)
                assert_same(vx.read_image(out), want, f"vx warp {kind} border {border} interp {interp:#x}")
                gg = ctypes.c_void_p(g)
                L.vxReleaseGraph(ctypes.byref(gg))
        # remap: the table set through vxCopyRemapPatch, converted on the GPU
        mp = synth.barrel_map(w, h, 0.12)
        rm = L.vxCreateRemap(ctx, w, h, w, h)
        rect = vx.Rect(0, 0, w, h)
        assert L.vxCopyRemapPatch(rm, ctypes.byref(rect), w * 8, mp.ctypes.data_as(ctypes.c_void_p), vx.VX_TYPE_COORDINATES2DF,
                                  vx.VX_WRITE_ONLY, vx.VX_MEMORY_TYPE_HOST) == 0
        assert L.vxuRemap(ctx, a, rm, vx.VX_INTERPOLATION_BILINEAR, out) == 0
        assert_same(vx.read_image(out), orc.remap(src, mp), "vxuRemap")
        # a changed table is picked up by the next run
        mp2 = synth.barrel_map(w, h, -0.1)
        assert L.vxCopyRemapPatch(rm, ctypes.byref(rect), w * 8, mp2.ctypes.data_as(ctypes.c_void_p), vx.VX_TYPE_COORDINATES2DF,
                                  vx.VX_WRITE_ONLY, vx.VX_MEMORY_TYPE_HOST) == 0
        assert L.vxuRemap(ctx, a, rm, vx.VX_INTERPOLATION_BILINEAR, out) == 0
        assert_same(vx.read_image(out), orc.remap(src, mp2), "vxuRemap after table update")
        assert L.vxuRemap(ctx, a, rm, vx.VX_INTERPOLATION_NEAREST_NEIGHBOR, out) == 0
        assert_same(vx.read_image(out), orc.remap_nearest(src, mp2)[0], "vxuRemap nearest")
    finally:
        h_ = ctypes.c_void_p(ctx)
        L.vxReleaseContext(ctypes.byref(h_))


@pytest.mark.parametrize("w,h", [(333, 77), (640, 480), (37, 29), (1920, 1080)])
def test_remap_rgb_is_three_u8_remaps(gpu, orc, w, h):
    """the stitch path's packed-RGB remap equals the U8 bilinear remap of each channel"""
    rgb = np.stack([synth.scene_u8(w, h, 60 + c) for c in range(3)], -1)
    s = Plane.from_array(rgb, "RGB")
    for k1 in (-0.2, 0.15):
        mp = synth.barrel_map(w, h, k1)
        mp[::5, ::3] = [-2.0, 1.0]
        fmap = torch.from_numpy(np.ascontiguousarray(mp)).cuda()
        fixed = torch.zeros((h, w * 4), dtype=torch.uint8, device="cuda")
        call("Remap_FixedMap_F32", w, h, ctypes.c_void_p(fixed.data_ptr()), w * 4, ctypes.c_void_p(fmap.data_ptr()), w * 8, w, h)
        for border in (None, 90):
            d = Plane(w, h, "RGB")
            if border is None:
                call("Remap_U24_U24_Bilinear", w, h, d, d.stride, w, h, s, s.stride, ctypes.c_void_p(fixed.data_ptr()), w * 4)
            else:
                call("Remap_U24_U24_Bilinear_Constant", w, h, d, d.stride, w, h, s, s.stride, ctypes.c_void_p(fixed.data_ptr()), w * 4,
                     ctypes.c_uint8(border))
            got = d.download()
            for c in range(3):
                want = orc.remap(np.ascontiguousarray(rgb[..., c]), mp, border is not None, border or 0)
                assert_same(got[..., c], want, f"RGB remap {w}x{h} k1 {k1} border {border} channel {c}")
