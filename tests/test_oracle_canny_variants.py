"""The oracle's Canny for gradient sizes 5 and 7 and the L2 norm (CannySobel_U16_U8_<g>x<g>_<norm>,
ago_haf_cpu_canny.cpp:145-427, 646-1056, with the gradient-7 suppression stage's in-place division of the
threshold by 4, ago_kernel_api.cpp:18033-18035), pinned against the reference engine (oracle/_ref) and
against closed forms."""
import numpy as np
import pytest

from mivisionx_amd import synth

THR = {3: (40, 110), 5: (400, 1300), 7: (1600, 5000)}


def _src(w, h, seed=7):
    return synth.scene_u8(w, h, seed) if w >= 16 and h >= 16 else synth.uniform_u8(w, h, seed)


@pytest.mark.parametrize("w,h", [(64, 48), (100, 60), (37, 29), (12, 10), (333, 77)])
@pytest.mark.parametrize("g", [3, 5, 7])
@pytest.mark.parametrize("l2", [0, 1])
def test_canny_variants_match_reference(orc, ref, w, h, g, l2):
    src = _src(w, h)
    lo, hi = THR[g] if not l2 else (THR[g][0] * 3 // 4, THR[g][1] * 3 // 4)
    want, after = ref.canny_ex(src, g, l2, lo, hi)
    eff = (lo // 4, hi // 4) if g == 7 else (lo, hi)
    assert after == eff
    got = orc.canny_ex(src, g, l2, *eff)
    assert np.array_equal(got, want), int((got != want).sum())


def test_gradient7_divides_the_threshold_every_run(orc, ref):
    src = _src(100, 60)
    want, after = ref.canny_ex(src, 7, 0, 1600, 5000, reps=2)
    assert after == (100, 312)
    assert np.array_equal(orc.canny_ex(src, 7, 0, 100, 312), want)


@pytest.mark.parametrize("g,l2", [(5, 0), (7, 0), (3, 1), (5, 1), (7, 1)])
def test_canny_sobel_closed_forms(orc, g, l2):
    """a constant image has zero gradient in every column the SSE passes reach from inside (sector 2 for
    (0, 0)); an x-ramp p = x gives Gx = (sum of smoothing) * (sum of k * derivative), Gy = 0"""
    w, h, R = 64, 24, g // 2
    c = np.full((h, w), 90, np.uint8)
    mag = orc.canny_sobel_ex(c, g, l2)
    inner = mag[R:h - R, R + 1:(w & ~15) - R - 1]
    assert ((inner >> 2) == 0).all() and ((inner & 3) == 2).all()
    ramp = np.tile(np.arange(w, dtype=np.uint8), (h, 1))
    mag = orc.canny_sobel_ex(ramp, g, l2).astype(np.int64)
    sm = {3: 4, 5: 16, 7: 64}[g]
    dsum = {3: 2, 5: 8, 7: 32}[g]          # sum over k of k * derivative tap
    gx = sm * dsum // (4 if g == 7 else 1)  # the 7x7 SSE columns carry the two >> 2 of its passes
    inner = mag[R:h - R, R + 1:(w & ~15) - R - 1]
    assert ((inner >> 2) == min(gx, 0x3fff if (g == 7 and not l2) else 1 << 30)).all()
    assert ((inner & 3) == 0).all()
