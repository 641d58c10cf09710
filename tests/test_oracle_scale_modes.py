"""The oracle's ScaleImage forms beside bilinear/UNDEFINED -- nearest neighbour, bilinear with REPLICATE and
CONSTANT borders, and area (nearest unless both dimensions shrink, exact 2x) -- pinned against the
reference's own CPU engine (oracle/_ref: vxScaleImageNode with the node border attribute, i.e.
ago_drama_divide.cpp:563-616 choosing ago_haf_cpu_geometric.cpp's kernels after the rewrite of
ago_drama_remove.cpp:1270-1311, and the AMD area kernel run by name) and against closed forms that
need no reference."""
import numpy as np
import pytest

from mivisionx_amd import synth
from pyoracle import BORDER_CONSTANT, BORDER_REPLICATE, BORDER_UNDEFINED, INTERP_AREA, INTERP_BILINEAR, INTERP_NEAREST

SIZES = [(200, 120, 100, 60), (200, 120, 97, 61), (64, 48, 160, 100), (37, 29, 90, 71), (640, 480, 1280, 960),
         (100, 80, 33, 17), (331, 211, 128, 96), (16, 16, 48, 40), (60, 40, 60, 20), (24, 8, 5, 3)]
MODES = [(BORDER_UNDEFINED, 0), (BORDER_REPLICATE, 0), (BORDER_CONSTANT, 0), (BORDER_CONSTANT, 200)]


def _src(sw, sh):
    return synth.scene_u8(sw, sh, sw * 7 + sh) if sw >= 16 and sh >= 16 else synth.uniform_u8(sw, sh, sw + sh)


@pytest.mark.parametrize("sw,sh,dw,dh", SIZES)
@pytest.mark.parametrize("interp", [INTERP_NEAREST, INTERP_BILINEAR, INTERP_AREA])
def test_scale_modes_match_reference(orc, ref, sw, sh, dw, dh, interp):
    if interp == INTERP_AREA and sw > dw and sh > dh and (sw, sh) != (2 * dw, 2 * dh):
        pytest.skip("a shrinking area scale other than 2x has no defined reference result")
    src = _src(sw, sh)
    for bm, bv in MODES:
        a, b = orc.scale_mode(src, dw, dh, interp, bm, bv), ref.scale_mode(src, dw, dh, interp, bm, bv)
        assert np.array_equal(a, b), (sw, sh, dw, dh, hex(interp), hex(bm), bv, int((a != b).sum()))


@pytest.mark.parametrize("sw,sh", [(200, 120), (64, 48), (34, 22), (1920, 1080), (50, 10)])
def test_area_2x_matches_reference(orc, ref, sw, sh):
    src = _src(sw, sh)
    for bm, bv in MODES:
        a, b = orc.scale_mode(src, sw // 2, sh // 2, INTERP_AREA, bm, bv), ref.scale_mode(src, sw // 2, sh // 2, INTERP_AREA, bm, bv)
        assert np.array_equal(a, b), (sw, sh, hex(bm))


@pytest.mark.parametrize("sw,sh,dw,dh", [(200, 120, 200, 120), (37, 29, 90, 71), (60, 40, 60, 20), (60, 40, 30, 40),
                                         (60, 40, 120, 20), (200, 120, 100, 60), (34, 22, 17, 11)])
def test_amd_area_kernel_matches_reference(orc, ref, sw, sh, dw, dh):
    """com.amd.openvx.ScaleImage_U8_U8_Area run by name: a copy at the same size, nearest when a
    dimension does not shrink, the 2x2 average at 2x"""
    src = _src(sw, sh)
    assert np.array_equal(orc.scale_area(src, dw, dh), ref.amd_scale_area(src, dw, dh))


def test_area_routing_closed_forms(orc):
    src = synth.scene_u8(64, 48, 9)
    assert np.array_equal(orc.scale_area(src, 64, 48), src)
    for dw, dh in ((160, 100), (64, 20), (30, 48)):
        assert np.array_equal(orc.scale_area(src, dw, dh), orc.scale_mode(src, dw, dh, INTERP_NEAREST))
    got = orc.scale_area(src, 32, 24)
    # row y averages source rows (int)(2y - 0.5) = 2y - 1 and 2y (row 0: rows 0 and 1)
    shifted = np.vstack([src[:2], src[1:-1]]).astype(np.int32)
    want = (shifted[0::2, 0::2] + shifted[0::2, 1::2] + shifted[1::2, 0::2] + shifted[1::2, 1::2] + 2) >> 2
    assert np.array_equal(got, want.astype(np.uint8))


@pytest.mark.parametrize("k", [2, 3, 4, 5])
def test_nearest_integer_downscale_closed_form(orc, k):
    """an exact downscale by k samples source (k x + k//2, k y + k//2): offset k/2, step k, truncated"""
    dw, dh = 37, 23
    src = synth.uniform_u8(dw * k, dh * k, k)
    got = orc.scale_mode(src, dw, dh, INTERP_NEAREST)
    want = src[k // 2::k, k // 2::k][:dh, :dw]
    assert np.array_equal(got, want)


@pytest.mark.parametrize("sw,sh,dw,dh", [(37, 29, 90, 71), (200, 120, 97, 61), (16, 16, 48, 40)])
def test_constant_image_stays_constant(orc, sw, sh, dw, dh):
    """weights of both roundings sum to 256 per axis: a constant image scales to itself (replicate, and
    constant border with the same value); area averages of a constant are the constant"""
    src = np.full((sh, sw), 173, np.uint8)
    for bm in (BORDER_REPLICATE, BORDER_CONSTANT):
        assert (orc.scale_mode(src, dw, dh, INTERP_BILINEAR, bm, 173) == 173).all(), hex(bm)
    assert (orc.scale_mode(src, dw, dh, INTERP_NEAREST) == 173).all()
    big = np.full((96, 160), 91, np.uint8)
    assert (orc.scale_area(big, 80, 48) == 91).all()


def test_constant_border_enters_only_at_edges(orc):
    """with a CONSTANT border the border value reaches the output only through the first and last
    source rows/columns: changing it leaves every output fed by interior rows and columns alone"""
    src = synth.scene_u8(64, 48, 3)
    a = orc.scale_mode(src, 160, 120, INTERP_BILINEAR, BORDER_CONSTANT, 0)
    b = orc.scale_mode(src, 160, 120, INTERP_BILINEAR, BORDER_CONSTANT, 255)
    diff = np.argwhere(a != b)
    assert len(diff) > 0
    assert (diff[:, 0] < 3).any() or (diff[:, 0] > 116).any()
    interior = (diff[:, 0] > 2) & (diff[:, 0] < 117) & (diff[:, 1] > 2) & (diff[:, 1] < 157)
    assert not interior.any()


def test_area_undefined_ratios_are_refused(orc):
    with pytest.raises(ValueError):
        orc.scale_area(synth.uniform_u8(90, 60, 1), 30, 20)
    with pytest.raises(ValueError):
        orc.scale_area(synth.uniform_u8(12, 8, 1), 6, 4)   # 2x but narrower than 8


@pytest.mark.parametrize("w,h", [(640, 480), (131, 67), (37, 29), (16, 16), (5, 5), (33, 4), (2, 3), (17, 9), (48, 3)])
@pytest.mark.parametrize("ksize", [1, 3, 5])
def test_halfscale_gaussian_matches_reference(orc, ref, w, h, ksize):
    """halfscale_gaussian kernel sizes 1 (nearest), 3 (ScaleGaussianHalf 3x3), 5 (5x5)"""
    if ksize == 5 and (w < 5 or h < 5):
        pytest.skip("ScaleGaussianHalf 5x5 needs a 5x5 source")
    src = _src(w, h)
    assert np.array_equal(orc.halfscale(src, ksize), ref.halfscale(src, ksize))


def test_gaussian_half_3x3_closed_forms(orc):
    """constant image: interior rows constant, column 0 at 3/4 (its left neighbour is the zero word);
    impulse at an even (row, column) => the [1 2 1]^2 / 16 kernel on the half grid, truncated"""
    c = np.full((40, 64), 160, np.uint8)
    got = orc.halfscale(c, 3)
    assert (got[0] == 0).all() and (got[-1] == 0).all()
    assert (got[1:-1, 1:] == 160).all() and (got[1:-1, 0] == 120).all()
    imp = np.zeros((40, 64), np.uint8)
    imp[20, 30] = 255
    got = orc.halfscale(imp, 3).astype(int)
    assert got[10, 15] == (4 * 255) >> 4
    imp[20, 30] = 0
    imp[21, 31] = 255   # odd row and column: reaches the four half-grid neighbours with weight 1
    got = orc.halfscale(imp, 3).astype(int)
    assert got[10, 15] == 255 >> 4 and got[11, 16] == 255 >> 4 and got[10, 16] == 255 >> 4 and got[11, 15] == 255 >> 4
    assert got.sum() == 4 * (255 >> 4)
