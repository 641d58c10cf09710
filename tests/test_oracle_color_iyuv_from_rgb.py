"""The oracle's RGB -> IYUV colour conversion, all three planes (HafCpu_ColorConvert_IYUV_RGB,
ago_haf_cpu_color_convert.cpp:2492-2705), pinned against the reference engine (a color_convert node with a
non-virtual IYUV output in oracle/_ref) and against closed forms."""
import numpy as np
import pytest


def _rgb(w, h, seed):
    return np.random.default_rng(seed).integers(0, 256, (h, w, 3), dtype=np.uint8)


@pytest.mark.parametrize("w,h", [(64, 48), (38, 30), (6, 4), (2, 2), (130, 10), (1280, 720)])
def test_iyuv_from_rgb_matches_reference(orc, ref, w, h):
    """body columns (below w & ~3) and the two-column scalar tail when w % 4 == 2"""
    rgb = _rgb(w, h, w * 7 + h)
    for a, b, name in zip(orc.color_iyuv_from_rgb(rgb), ref.color_iyuv_from_rgb(rgb), "YUV"):
        assert np.array_equal(a, b), name


def test_iyuv_from_rgb_extremes_match_reference(orc, ref):
    """saturated and black pixels, pure primaries: the clamps and the largest chroma excursions"""
    w, h = 38, 8
    vals = np.array([0, 255], np.uint8)
    rgb = vals[np.indices((h, w, 3)).sum(0) % 2]
    rgb[::3] = [255, 0, 0]
    rgb[1::5, ::2] = [0, 0, 255]
    for a, b, name in zip(orc.color_iyuv_from_rgb(rgb), ref.color_iyuv_from_rgb(rgb), "YUV"):
        assert np.array_equal(a, b), name


def test_iyuv_from_rgb_closed_forms(orc):
    """the luma plane equals ColorConvert Y (the pinned single-plane kernel) on body columns; a flat
    colour gives flat chroma equal to its single-pixel value; grey gives chroma within 1 of 128"""
    w, h = 64, 8
    rgb = _rgb(w, h, 3)
    y, u, v = orc.color_iyuv_from_rgb(rgb)
    assert np.array_equal(y, orc.color_y(rgb))
    for g in (0, 17, 128, 255):
        flat = np.full((4, 8, 3), g, np.uint8)
        _, u, v = orc.color_iyuv_from_rgb(flat)
        assert (np.abs(u.astype(int) - 128) <= 1).all() and (np.abs(v.astype(int) - 128) <= 1).all()
        assert (u == u[0, 0]).all() and (v == v[0, 0]).all()
