"""Pins the oracle's 3x3 filters, Sobel, Magnitude and Phase (CPU only).

Three independent checks: the golden hashes written from the reference CPU engine, the live
reference engine (when oracle/_ref is built), and closed forms that a dropped term, a wrong sign
or a shifted index would break.
"""
import numpy as np
import pytest

from mivisionx_amd import synth
from util import assert_same, sha

FILTERS = ("gaussian", "box", "median")


def test_inputs_match_golden(golden):
    for c in golden["cases"]:
        assert sha(synth.uniform_u8(c["w"], c["h"], c["seed"])) == c["input_sha256"], (c["w"], c["h"], c["seed"])


def test_oracle_matches_golden(orc, golden):
    for c in golden["cases"]:
        src = synth.uniform_u8(c["w"], c["h"], c["seed"])
        tag = f'{c["w"]}x{c["h"]} seed {c["seed"]}'
        gx, gy, mag, ph = orc.chain1(src)
        for k, v in dict(gx=gx, gy=gy, mag=mag, phase=ph).items():
            assert sha(v) == c["chain1"][k]["sha256"], f"chain1.{k} {tag}"
        for which, name in enumerate(FILTERS):
            assert sha(orc.filter3x3(which, src)) == c[name]["out"]["sha256"], f"{name} {tag}"
        sx, sy = orc.sobel(src)
        assert sha(sx) == c["sobel"]["gx"]["sha256"], f"sobel.gx {tag}"
        assert sha(sy) == c["sobel"]["gy"]["sha256"], f"sobel.gy {tag}"


def _grid(lim):
    g = np.arange(-lim, lim + 1, dtype=np.int16)
    gx = np.ascontiguousarray(np.broadcast_to(g[None, :], (g.size, g.size)))
    gy = np.ascontiguousarray(np.broadcast_to(g[:, None], (g.size, g.size)))
    return gx, gy


def test_magnitude_phase_grid_matches_golden(orc, golden):
    gg = golden["gradient_grid"]
    gx, gy = _grid(gg["lim"])
    mag = orc.magnitude(gx, gy)
    ph = orc.phase(gx, gy)
    assert sha(mag) == gg["magnitude"]["sha256"]
    assert sha(ph) == gg["phase"]["sha256"]
    for key, (m, p) in gg["samples"].items():
        x, y = map(int, key.split(","))
        assert (int(mag[y + 1020, x + 1020]), int(ph[y + 1020, x + 1020])) == (m, p), key


@pytest.mark.parametrize("w,h", [(48, 7), (100, 33), (17, 4), (1, 3), (2, 5), (640, 480)])
@pytest.mark.parametrize("seed", [3, 11])
def test_oracle_matches_live_reference(orc, ref, w, h, seed):
    src = synth.uniform_u8(w, h, seed)
    for a, b, k in zip(orc.chain1(src), ref.chain1(src), ("gx", "gy", "mag", "phase")):
        assert_same(a, b, f"chain1.{k} {w}x{h}")
    for which, name in enumerate(FILTERS):
        assert_same(orc.filter3x3(which, src), ref.filter3x3(which, src), f"{name} {w}x{h}")
    for a, b, k in zip(orc.sobel(src), ref.sobel(src), ("gx", "gy")):
        assert_same(a, b, f"sobel.{k} {w}x{h}")


def test_oracle_magnitude_phase_live_reference_s16_extremes(orc, ref):
    # full S16 range, including the saturating and wrapping corners of Magnitude's two rules
    rng = np.random.default_rng(5)
    gx = rng.integers(-32768, 32768, (9, 37), dtype=np.int16)
    gy = rng.integers(-32768, 32768, (9, 37), dtype=np.int16)
    gx[0, :4] = [-32768, 32767, 0, -32768]
    gy[0, :4] = [-32768, 32767, -32768, 0]
    m, p = ref.magnitude_phase(gx, gy)
    assert_same(orc.magnitude(gx, gy), m, "magnitude")
    assert_same(orc.phase(gx, gy), p, "phase")


# ---------------- closed forms (independent of the reference) ----------------

@pytest.mark.parametrize("c", [0, 1, 77, 255])
def test_constant_image_filters(orc, c):
    w, h = 32, 9   # stride == width, so the wrap columns read real pixels except at the two guards
    src = np.full((h, w), c, np.uint8)
    for which in range(3):
        out = orc.filter3x3(which, src)
        exp = np.full((h, w), c, np.uint8)
        exp[0, :] = 0
        exp[h - 1, :] = 0
        # (x=0, y=1) misses its top-left neighbour and (x=w-1, y=h-2) its bottom-right one: those
        # two reads fall on the zero guard bytes before/after the buffer
        if which == 0:
            exp[1, 0] = exp[h - 2, w - 1] = (15 * c) >> 4
        elif which == 1:
            exp[1, 0] = exp[h - 2, w - 1] = (8 * c) // 9
        else:
            exp[1, 0] = exp[h - 2, w - 1] = c
        assert_same(out, exp, f"filter {which} c={c}")
    gx, gy = orc.sobel(src)
    inner = (slice(2, h - 2), slice(1, w - 1))
    assert not gx[inner].any() and not gy[inner].any()


def test_sobel_ramp(orc):
    w, h, a, b = 48, 12, 3, 5
    y, x = np.mgrid[0:h, 0:w]
    src = (a * x + b * y).astype(np.uint8)   # max 3*47 + 5*11 = 196: no wrap
    gx, gy = orc.sobel(src)
    assert (gx[1:h - 1, 1:w - 1] == 8 * a).all()
    assert (gy[1:h - 1, 1:w - 1] == 8 * b).all()
    assert not gx[0].any() and not gx[h - 1].any()


def test_gaussian_impulse_is_the_binomial_kernel(orc):
    src = np.zeros((9, 32), np.uint8)
    src[4, 10] = 255
    out = orc.filter3x3(0, src).astype(int)
    k = np.array([[1, 2, 1], [2, 4, 2], [1, 2, 1]])
    assert (out[3:6, 9:12] == (255 * k) >> 4).all()
    out[3:6, 9:12] = 0
    assert not out.any()


def test_box_mulhi_identity():
    s = np.arange(0, 9 * 255 + 1, dtype=np.int64)
    assert ((s * 7282) >> 16 == s // 9).all()


def test_magnitude_closed_form(orc):
    gx, gy = _grid(1020)
    mag = orc.magnitude(gx, gy).astype(np.int64)
    s = gx.astype(np.int64) ** 2 + gy.astype(np.int64) ** 2
    k = np.floor(np.sqrt(s.astype(np.float64))).astype(np.int64)
    k -= (k * k > s)
    k += ((k + 1) * (k + 1) <= s)
    exp = np.where(s - k * k > k, k + 1, k)
    assert (mag == exp).all()


def test_phase_axes_and_accuracy(orc):
    px = orc.lib.oracle_phase_px
    assert [px(5, 0), px(0, 5), px(-5, 0), px(0, -5), px(0, 0)] == [0, 64, 128, 192, 0]
    gx, gy = _grid(300)
    ph = orc.phase(gx, gy).astype(np.int64)
    exact = np.degrees(np.arctan2(gy.astype(np.float64), gx.astype(np.float64))) % 360.0
    ideal = np.floor(exact * 128.0 / 180.0 + 0.5).astype(np.int64) & 0xFF
    diff = np.minimum((ph - ideal) % 256, (ideal - ph) % 256)
    assert diff.max() <= 1
