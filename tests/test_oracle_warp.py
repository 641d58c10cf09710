"""Pins the oracle's WarpAffine, WarpPerspective and Remap restatements (CPU only): golden hashes of
the reference engine's outputs, the live reference engine when oracle/_ref is built, and closed
forms that do not reuse the oracle's formulas (identity, integer shifts, integer remap tables)."""
import json
import os

import numpy as np
import pytest

from mivisionx_amd import synth
from util import assert_same, sha

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def gw():
    with open(os.path.join(ROOT, "tests", "golden", "vision_golden_warp.json")) as f:
        return json.load(f)


def _src(w, h, g):
    src = synth.scene_u8(w, h, g["scene_seed"])
    return src


def test_warps_match_golden(orc, gw):
    n = 0
    for c in gw["cases"]:
        if "affine" not in c:
            continue
        src = _src(c["w"], c["h"], gw)
        assert sha(src) == c["src_sha256"]
        const, border = c["border"] is not None, c["border"] or 0
        got, unp = orc.warp_affine(src, c["dw"], c["dh"], np.array(c["affine"], np.float32), const, border)
        got[unp] = 0
        assert int(unp.sum()) == c["affine_unpinned"] and sha(got) == c["affine_sha256"], ("affine", c["w"], c["h"], c["seed"], c["border"])
        got, unp = orc.warp_perspective(src, c["dw"], c["dh"], np.array(c["perspective"], np.float32), border)
        got[unp] = 0
        assert int(unp.sum()) == c["perspective_unpinned"] and sha(got) == c["perspective_sha256"], \
            ("perspective", c["w"], c["h"], c["seed"], c["border"])
        n += 1
    assert n >= 40


def test_remap_matches_golden(orc, gw):
    n = 0
    for c in gw["cases"]:
        if "k1" not in c:
            continue
        src = _src(c["w"], c["h"], gw)
        mp = synth.barrel_map(c["w"], c["h"], c["k1"])
        assert sha(mp) == c["map_sha256"]
        got = orc.remap(src, mp, c["border"] is not None, c["border"] or 0)
        assert sha(got) == c["remap_sha256"], (c["w"], c["h"], c["k1"], c["border"])
        n += 1
    assert n >= 16


@pytest.mark.parametrize("w,h", [(97, 41), (256, 130)])
def test_oracle_matches_live_reference(orc, ref, w, h):
    src = synth.scene_u8(w, h, 30)
    for seed in range(3, 7):
        aff, per = synth.warp_matrices(w, h, seed)
        for border in (None, 5):
            got, unp = orc.warp_affine(src, w, h, aff, border is not None, border or 0)
            want = ref.warp(0, src, w, h, aff, border is not None, border or 0)
            assert_same(np.where(unp, 0, got), np.where(unp, 0, want), f"affine seed {seed}")
            got, unp = orc.warp_perspective(src, w, h, per, border or 0)
            want = ref.warp(1, src, w, h, per, border is not None, border or 0)
            assert_same(np.where(unp, 0, got), np.where(unp, 0, want), f"perspective seed {seed}")
    mp = synth.barrel_map(w, h, 0.1)
    mp[::7, ::5] = -3.0          # out-of-source points (the 0xffff entries)
    for border in (None, 31):
        assert_same(orc.remap(src, mp, border is not None, border or 0), ref.remap(src, mp, border is not None, border or 0), "remap")


def test_identity_and_integer_shift_closed_forms(orc):
    w, h = 50, 20
    src = synth.uniform_u8(w, h, 31)
    ident = np.array([1, 0, 0, 1, 0, 0], np.float32)
    got, _ = orc.warp_affine(src, w, h, ident, True, 7)
    assert_same(got, src, "affine identity")
    got, _ = orc.warp_perspective(src, w, h, np.array([1, 0, 0, 0, 1, 0, 0, 0, 1], np.float32), 7)
    assert_same(got, src, "perspective identity")
    # x_src = x + 3, y_src = y + 2: a shifted copy; outside the source the CONSTANT border
    got, _ = orc.warp_affine(src, w, h, np.array([1, 0, 0, 1, 3, 2], np.float32), True, 7)
    want = np.full((h, w), 7, np.uint8)
    want[:h - 2, :w - 3] = src[2:, 3:]
    assert_same(got, want, "affine integer shift")
    # an integer remap table (fraction 0) copies the source pixel
    yy, xx = np.mgrid[0:h, 0:w]
    mp = np.stack([np.minimum(xx + 1, w - 2), np.minimum(yy, h - 2)], -1).astype(np.float32)
    got = orc.remap(src, mp, True, 0)
    assert_same(got, src[np.minimum(yy, h - 2), np.minimum(xx + 1, w - 2)], "remap integer table")


def test_remap_fixed_point_table(orc):
    mp = np.array([[[0.0, 0.0], [1.06, 2.5], [-0.01, 3.0], [9.0, 1.0], [8.99, 8.99], [3.3125, 0.0624]]], np.float32)
    fx = orc.remap_fixed(mp, 10, 10)
    # round(v * 8) with +0.5, and (0xffff, 0xffff) for sx < 0, sy < 0, sx >= W-1 or sy >= H-1
    assert fx[0].tolist() == [[0, 0], [8, 20], [0xffff, 0xffff], [0xffff, 0xffff], [72, 72], [27, 0]]


INTERIOR = (np.array([0.9, 0.05, -0.04, 0.92, 3.3, 2.7], np.float32),
            np.array([0.9, 0.05, 0.0001, -0.04, 0.92, 0.0002, 3.3, 2.7, 1.0], np.float32))


def _cmp_rows(got, want, dw, what, unp=None):
    """the first ceil4(dw) bytes of every row (the pixels and the padding the 4-pixel stores write)"""
    q = (dw + 3) & ~3
    diff = got[:, :q] != want[:, :q]
    if unp is not None:
        diff[:, :dw] &= ~unp
    assert not diff.any(), f"{what}: {int(diff.sum())} differ, first {np.argwhere(diff)[:4].tolist()}"


@pytest.mark.parametrize("w,h", [(333, 77), (37, 29), (44, 30)])
def test_padding_and_nearest_match_live_reference(orc, ref, w, h):
    """row padding of the bilinear warps/remap and every nearest-neighbour form, against the reference
    engine's own buffers (read with their padding); nearest affine only at widths that are a multiple of
    4, where the reference does not fault"""
    from pyoracle import INTERP_BILINEAR, INTERP_NEAREST
    src = synth.scene_u8(w, h, 40)
    for seed in (1, 2, 0):
        aff, per = synth.warp_matrices(w, h, seed) if seed else INTERIOR
        for dw, dh in ((w, h), (w + 5, max(3, h - 3)), (w + 4, h + 2)):
            for border in (None, 77):
                cst, bv = border is not None, border or 0
                tag = f"{w}x{h}->{dw}x{dh} seed {seed} border {border}"
                got, unp = orc.warp_affine(src, dw, dh, aff, cst, bv, padded=True)
                _cmp_rows(got, ref.warp_padded(0, INTERP_BILINEAR, src, dw, dh, aff, cst, bv), dw, "affine bilinear " + tag, unp)
                got, unp = orc.warp_perspective(src, dw, dh, per, bv, padded=True)
                _cmp_rows(got, ref.warp_padded(1, INTERP_BILINEAR, src, dw, dh, per, cst, bv), dw, "perspective bilinear " + tag, unp)
                if dw % 4 == 0:
                    got = orc.warp_affine_nearest(src, dw, dh, aff, border, padded=True)
                    _cmp_rows(got, ref.warp_padded(0, INTERP_NEAREST, src, dw, dh, aff, cst, bv), dw, "affine nearest " + tag)
                got = orc.warp_perspective_nearest(src, dw, dh, per, border, padded=True)
                _cmp_rows(got, ref.warp_padded(1, INTERP_NEAREST, src, dw, dh, per, cst, bv), dw, "perspective nearest " + tag)
    for k in (0.12, -0.1):
        mp = synth.barrel_map(w, h, k)
        mp[::7, ::5] = -3.0
        for border in (None, 5, 20):
            cst, bv = border is not None, border or 0
            got = orc.remap(src, mp, cst, bv, padded=True)
            _cmp_rows(got, ref.remap_padded(INTERP_BILINEAR, src, mp, cst, bv), w, f"remap bilinear k {k} border {border}")
            got, unp = orc.remap_nearest(src, mp, cst, bv, padded=True)
            want = ref.remap_padded(INTERP_NEAREST, src, mp, cst, bv)
            _cmp_rows(got, want, w, f"remap nearest k {k} border {border}", unp)
            assert not want[:, w:].any() and not got[:, w:].any()   # the nearest remap writes no padding


def test_nearest_closed_forms(orc):
    w, h = 52, 20
    src = synth.uniform_u8(w, h, 32)
    ident = np.array([1, 0, 0, 1, 0, 0], np.float32)
    assert_same(orc.warp_affine_nearest(src, w, h, ident, 7), src, "affine nearest identity")
    assert_same(orc.warp_perspective_nearest(src, w, h, np.array([1, 0, 0, 0, 1, 0, 0, 0, 1], np.float32), 7), src,
                "perspective nearest identity")
    # x_src = x + 3.6, y_src = y + 2.2: truncation picks (x + 3, y + 2); CONSTANT border outside
    got = orc.warp_affine_nearest(src, w, h, np.array([1, 0, 0, 1, 3.6, 2.2], np.float32), 7)
    want = np.full((h, w), 7, np.uint8)
    want[:h - 2, :w - 4] = src[2:, 3:w - 1]
    want[:h - 2, w - 4] = src[2:, w - 1]
    assert_same(got, want, "affine nearest shift")
    # padding pixels of a warp repeat column 0
    got = orc.warp_perspective_nearest(src, 50, h, np.array([1, 0, 0, 0, 1, 0, 1.5, 0.5, 1], np.float32), None, padded=True)
    assert (got[:, 50] == got[:, 0]).all() and (got[:, 51] == got[:, 0]).all()
    # remap: fraction > 3/8 rounds up in the 4-pixel body and in the tail alike; out-of-source points
    # read coordinate 1 (body) or 0 (tail) under UNDEFINED
    yy, xx = np.mgrid[0:h, 0:w - 1]
    mp = np.stack([np.minimum(xx + 0.5, w - 3), np.minimum(yy + 0.25, h - 3)], -1).astype(np.float32)
    got, _ = orc.remap_nearest(src, mp)
    assert_same(got, src[np.minimum(yy, h - 3), np.minimum(xx + 1, w - 3)], "remap nearest rounding")
    mp[:, :] = -1.0
    got, _ = orc.remap_nearest(src, mp)
    body = (w - 1) & ~3
    assert (got[:, :body] == src[1, 1]).all() and (got[:, body:] == src[0, 0]).all()
    got, _ = orc.remap_nearest(src, mp, True, 9)
    assert (got[:, :body] == 9).all() and (got[:, body:] == src[9, 9]).all()   # the tail reads at (border, border)
