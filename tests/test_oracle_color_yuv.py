"""The oracle's YUV -> RGB colour conversion (ColorConvert_RGB_IYUV / _RGB_NV12,
ago_haf_cpu_color_convert.cpp:472-768), pinned against the reference engine (color_convert nodes in
oracle/_ref) and against closed forms."""
import numpy as np
import pytest


def _planes(w, h, seed):
    rng = np.random.default_rng(seed)
    return (rng.integers(0, 256, (h, w), dtype=np.uint8), rng.integers(0, 256, (h // 2, w // 2), dtype=np.uint8),
            rng.integers(0, 256, (h // 2, w // 2), dtype=np.uint8))


@pytest.mark.parametrize("w,h", [(64, 48), (100, 60), (38, 30), (1280, 720), (16, 2), (18, 10)])
def test_rgb_from_yuv_matches_reference(orc, ref, w, h):
    y, u, v = _planes(w, h, w + h)
    assert np.array_equal(orc.color_rgb_yuv(y, u, v), ref.color_rgb_yuv(y, u, v))
    uv = np.stack([u, v], -1).reshape(h // 2, w)
    assert np.array_equal(orc.color_rgb_yuv(y, uv), ref.color_rgb_yuv(y, uv))


def test_rgb_from_yuv_closed_forms(orc):
    """grey chroma (128) gives R = G = B = Y; IYUV and NV12 carrying the same samples agree; a pure
    +v chroma raises only R and lowers G"""
    w, h = 32, 8
    y, u, v = _planes(w, h, 1)
    grey = np.full((h // 2, w // 2), 128, np.uint8)
    rgb = orc.color_rgb_yuv(y, grey, grey)
    assert (rgb == y[..., None]).all()
    uv = np.stack([u, v], -1).reshape(h // 2, w)
    assert np.array_equal(orc.color_rgb_yuv(y, u, v), orc.color_rgb_yuv(y, uv))
    flat = np.full((h, w), 100, np.uint8)
    rgb = orc.color_rgb_yuv(flat, grey, np.full_like(grey, 178)).astype(int)   # v = 50
    assert (rgb[..., 0] == 100 + int(np.float32(50) * np.float32(1.5748))).all()
    assert (rgb[..., 2] == 100).all() and (rgb[..., 1] < 100).all()
