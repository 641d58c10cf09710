"""Pins the oracle's ColorConvert luma, ScaleImage, GaussianPyramid, Canny and Harris (CPU only):
golden hashes from the reference engine, the live reference engine when oracle/_ref is built, and
closed forms / invariants that do not reuse the oracle's formulas."""
import numpy as np
import pytest

from mivisionx_amd import synth
from pyoracle import KEYPOINT  # noqa: F401
from util import assert_same, sha


@pytest.fixture(scope="module")
def g2(request):
    import json
    import os
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    with open(os.path.join(root, "tests", "golden", "vision_golden_graph2.json")) as f:
        return json.load(f)


def test_nodes_match_golden(orc, g2):
    for c in g2["nodes"]:
        w, h = c["w"], c["h"]
        rgb = synth.scene_rgb(w, h, c["seed"])
        assert sha(rgb) == c["rgb_sha256"], (w, h)
        y = orc.color_y(rgb)
        assert sha(y) == c["color_y"], f"color {w}x{h}"
        for key, want in c["scale"].items():
            dw, dh = map(int, key.split("x"))
            assert sha(orc.scale(y, dw, dh)) == want, f"scale {w}x{h}->{key}"
        if "pyramid3" in c:
            assert [sha(l) for l in orc.pyramid(y, 3)] == c["pyramid3"], f"pyramid {w}x{h}"
        for key, want in c["canny"].items():
            lo, hi = map(int, key.split(","))
            assert sha(orc.canny(y, lo, hi)) == want, f"canny {key} {w}x{h}"
        for hc in c["harris"]:
            kp, n = orc.harris(y, hc["k"], hc["thr"], hc["dist"], hc["capacity"])
            assert n == hc["count"] and sha(kp) == hc["sha256"], f"harris {hc} {w}x{h}"


def test_graph2_4k_matches_golden(orc, g2):
    g = g2["graph2"]
    rgb = synth.scene_rgb(g["w"], g["h"], g["seed"])
    assert sha(rgb) == g["rgb_sha256"]
    r = orc.graph2(rgb, g["dw"], g["dh"], g["levels"], g["lower"], g["upper"], g["k"], g["thr"], g["dist"], g["capacity"])
    assert sha(r["scaled"]) == g["scaled"]
    assert [sha(l) for l in r["pyramid"]] == g["pyramid"]
    assert sha(r["edges"]) == g["edges"]
    assert r["count"] == g["count"] and sha(r["corners"]) == g["corners"]


@pytest.mark.parametrize("w,h", [(48, 30), (66, 34), (18, 12), (256, 64)])
def test_oracle_matches_live_reference(orc, ref, w, h):
    rgb = synth.scene_rgb(w, h, 9)
    y = ref.color_y(rgb)
    assert_same(orc.color_y(rgb), y, "color")
    for (dw, dh) in [(w // 2, h // 2), (w + 7, h + 3), (w // 3, h // 3)]:
        assert_same(orc.scale(y, dw, dh), ref.scale(y, dw, dh), f"scale {dw}x{dh}")
    for a, b in zip(orc.pyramid(y, 2), ref.pyramid(y, 2)):
        assert_same(a, b, "pyramid")
    for lo, hi in [(80, 100), (5, 20)]:
        assert_same(orc.canny(y, lo, hi), ref.canny(y, lo, hi), f"canny {lo},{hi}")
    noise = synth.uniform_u8(w, h, 4)
    assert_same(orc.canny(noise, 150, 250), ref.canny(noise, 150, 250), "canny noise")
    for (k, thr, d, cap) in [(0.04, 1e-6, 5.0, 100), (0.1, 1e-7, 3.0, 1000)]:
        a, na = orc.harris(noise, k, thr, d, cap)
        b, nb = ref.harris(noise, k, thr, d, cap)
        assert na == nb
        assert_same(a, b, f"harris {k},{thr},{d}")


# ---------------- closed forms / invariants ----------------

def test_color_primaries_and_greys(orc):
    rgb = np.zeros((2, 256, 3), np.uint8)
    rgb[0, :, :] = np.arange(256)[:, None]            # greys
    rgb[1, :3] = [[255, 0, 0], [0, 255, 0], [0, 0, 255]]
    y = orc.color_y(rgb)
    assert list(y[1, :3]) == [54, 182, 18]            # trunc(255 * 0.2126), (* 0.7152), (* 0.0722)
    assert (y[0] <= np.arange(256)).all() and (y[0] >= np.arange(256) - 1).all()


def test_scale_exact_half_is_two_stage_rounded_average(orc):
    src = synth.uniform_u8(64, 40, 8)
    out = orc.scale(src, 32, 20).astype(int)
    s = src.astype(int)
    t0 = (s[0::2, 0::2] + s[0::2, 1::2] + 1) >> 1
    t1 = (s[1::2, 0::2] + s[1::2, 1::2] + 1) >> 1
    assert (out == ((t0 + t1 + 1) >> 1)).all()


def test_pyramid_constant_and_impulse(orc):
    c = np.full((41, 64), 90, np.uint8)
    lvl = orc.gaussian_half(c, 32, 21)
    assert (lvl[2:19, 2:30] == 90).all() and not lvl[0].any() and not lvl[20].any()
    imp = np.zeros((41, 64), np.uint8)
    imp[20, 33] = 255                                   # odd height: rows 2y; even width: columns 2x+1
    lvl = orc.gaussian_half(imp, 32, 21).astype(int)
    k = np.array([1, 4, 6, 4, 1])
    assert lvl[10, 16] == (255 * 36) >> 8               # centre tap 6*6
    assert lvl[9, 16] == (255 * 6) >> 8                 # two source rows away: taps 1 (vertical) x 6
    assert lvl.sum() == sum((255 * k[j] * k[i]) >> 8 for j in (0, 2, 4) for i in (0, 2, 4))


def test_canny_trace_equals_connected_components(orc):
    from scipy import ndimage
    y = synth.scene_u8(160, 96, 12)
    pre, _ = orc.canny_supp_threshold(orc.canny_sobel(y), 30, 90)
    lab, n = ndimage.label(pre > 0, structure=np.ones((3, 3), int))
    strong_labels = np.unique(lab[pre == 255])
    want = np.where(np.isin(lab, strong_labels[strong_labels > 0]), 255, 0).astype(np.uint8)
    assert_same(orc.canny_trace(pre), want, "trace vs components")
    assert_same(orc.canny(y, 30, 90), want, "canny node")


def test_harris_pick_invariants(orc):
    y = synth.uniform_u8(300, 200, 21)
    d = 7.0
    kp, n, hvc = orc.harris(y, 0.04, 1e-7, d, 100000, with_score=True)
    assert n == len(kp) > 50
    keys = [(float(k["strength"]), int(k["y"]), int(k["x"])) for k in kp]
    assert keys == sorted(keys, reverse=True)
    xy = np.stack([kp["x"], kp["y"]], 1).astype(int)
    d2 = ((xy[:, None, :] - xy[None, :, :]) ** 2).sum(-1)
    np.fill_diagonal(d2, 10 ** 9)
    assert d2.min() >= int(np.ceil(d * d))
    for k in kp:   # every corner is a strict/weak local maximum of the score, as the NMS keeps it
        x, yy = int(k["x"]), int(k["y"])
        win = hvc[yy - 1:yy + 2, x - 1:x + 2]
        assert win[1, 1] == k["strength"] and win[1, 1] >= win[0].max() and win[1, 1] > win[2].max()
