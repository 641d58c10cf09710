"""The ColorConvert oracle (oracle/color.c, geometric.c) for every conversion the reference's CPU engine
runs: bit-exact against the reference engine (oracle/_ref) on random images whose widths cover each
kernel's SSE body and scalar tail, against the committed golden hashes (which the reference engine
wrote), and against closed forms that hold without it."""
import json
import os

import numpy as np
import pytest

from pyoracle import COLOR_CONVERSIONS, color_planes
from util import assert_same, sha

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# widths: below one SSE step, 16/32-pixel bodies with 2-, 6- and 14-pixel tails, exact multiples
SIZES = [(2, 2), (16, 4), (18, 6), (22, 2), (30, 8), (34, 4), (64, 10), (66, 6), (70, 2), (96, 4), (102, 6)]


def _planes(fmt, w, h, seed):
    rng = np.random.default_rng(seed)
    return [rng.integers(0, 256, shape, dtype=np.uint8) for shape in color_planes(fmt, w, h)]


@pytest.mark.parametrize("src,dst", COLOR_CONVERSIONS)
def test_color_convert_vs_reference_engine(orc, ref, src, dst):
    for k, (w, h) in enumerate(SIZES):
        planes = _planes(src, w, h, 1000 * k + w)
        want = ref.color_convert(src, dst, w, h, planes)
        got = orc.color_convert(src, dst, w, h, planes)
        for p, (g, r) in enumerate(zip(got, want)):
            assert_same(g, r, f"{src}->{dst} {w}x{h} plane {p}")


@pytest.mark.parametrize("src", ["RGB", "RGBX"])
@pytest.mark.parametrize("yuv", ["IYUV", "NV12", "YUV4"])
def test_luma_only_conversion_vs_reference_engine(orc, ref, src, yuv):
    """a virtual YUV output read only through channel_extract(Y): the reference keeps only the luma
    kernel (ColorConvert_Y_RGB / _Y_RGBX), whose rounding differs from the merged kernels'"""
    for k, (w, h) in enumerate(SIZES):
        planes = _planes(src, w, h, 77 + k)
        assert_same(orc.color_luma(src, w, h, planes), ref.color_luma(src, yuv, w, h, planes), f"{src}->{yuv} luma {w}x{h}")


def test_color_convert_golden(orc):
    """the reference engine's outputs as committed hashes (tests/golden/make_golden_color.py)"""
    with open(os.path.join(ROOT, "tests", "golden", "vision_golden_color.json")) as f:
        golden = json.load(f)
    assert len(golden["cases"]) >= len(COLOR_CONVERSIONS)
    for c in golden["cases"]:
        planes = _planes(c["src"], c["w"], c["h"], c["seed"])
        if c.get("luma_only"):
            got = [orc.color_luma(c["src"], c["w"], c["h"], planes)]
        else:
            got = orc.color_convert(c["src"], c["dst"], c["w"], c["h"], planes)
        assert [sha(g) for g in got] == c["sha"], f"{c['src']}->{c['dst']} {c['w']}x{c['h']}"


@pytest.mark.parametrize("src", ["UYVY", "YUYV", "IYUV", "NV12", "NV21"])
@pytest.mark.parametrize("dst", ["RGB", "RGBX"])
def test_grey_chroma_gives_grey(orc, src, dst):
    """U = V = 128: R = G = B = Y exactly, X = 255"""
    w, h = 38, 6
    planes = _planes(src, w, h, 5)
    if src in ("UYVY", "YUYV"):
        p = planes[0].reshape(h, w // 2, 4)
        p[..., [0, 2] if src == "UYVY" else [1, 3]] = 128
        y = p[..., [1, 3] if src == "UYVY" else [0, 2]].reshape(h, w)
    else:
        for c in planes[1:]:
            c[:] = 128
        y = planes[0]
    out = orc.color_convert(src, dst, w, h, planes)[0].reshape(h, w, 3 if dst == "RGB" else 4)
    for k in range(3):
        assert_same(out[..., k], y, f"{src}->{dst} channel {k}")
    if dst == "RGBX":
        assert (out[..., 3] == 255).all()


def test_round_trips(orc):
    """RGB -> RGBX -> RGB and IYUV -> NV12 -> IYUV are identities; NV21 -> IYUV swaps NV12's chroma;
    IYUV -> YUV4 and NV12 -> YUV4 replicate each chroma sample over its 2x2 block"""
    w, h = 34, 10
    rgb = _planes("RGB", w, h, 1)
    assert_same(orc.color_convert("RGBX", "RGB", w, h, orc.color_convert("RGB", "RGBX", w, h, rgb))[0], rgb[0], "RGB->RGBX->RGB")
    iyuv = _planes("IYUV", w, h, 2)
    back = orc.color_convert("NV12", "IYUV", w, h, orc.color_convert("IYUV", "NV12", w, h, iyuv))
    for a, b in zip(back, iyuv):
        assert_same(a, b, "IYUV->NV12->IYUV")
    nv = _planes("NV12", w, h, 3)
    a = orc.color_convert("NV12", "IYUV", w, h, nv)
    b = orc.color_convert("NV21", "IYUV", w, h, nv)
    assert_same(a[1], b[2], "NV21 U is NV12 V")
    assert_same(a[2], b[1], "NV21 V is NV12 U")
    y4 = orc.color_convert("IYUV", "YUV4", w, h, iyuv)
    assert_same(y4[1], np.repeat(np.repeat(iyuv[1], 2, 0), 2, 1), "IYUV->YUV4 U")
    assert_same(y4[2], np.repeat(np.repeat(iyuv[2], 2, 0), 2, 1), "IYUV->YUV4 V")
