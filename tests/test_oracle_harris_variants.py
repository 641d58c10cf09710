"""The oracle's Harris for gradient and block sizes 5 and 7 (HarrisSobel_HG3_U8_5x5/7x7,
HarrisScore_HVC_HG3_5x5/7x7, ago_haf_cpu_harris.cpp:333-802, with the norm of ago_kernel_api.cpp:17012 and
:17083), pinned against the reference engine (oracle/_ref) through the whole corner node and against
closed forms."""
import numpy as np
import pytest

from mivisionx_amd import synth


@pytest.mark.parametrize("w,h", [(64, 48), (101, 61), (37, 29), (333, 77)])
@pytest.mark.parametrize("g", [3, 5, 7])
@pytest.mark.parametrize("b", [3, 5, 7])
def test_harris_variants_match_reference(orc, ref, w, h, g, b):
    src = synth.scene_u8(w, h, 11)
    thr = 1e-4 if g == 3 else 1e-5
    want, wc = ref.harris_ex(src, g, b, 0.04, thr, 3.0, 2000)
    got, gc = orc.harris_ex(src, g, b, 0.04, thr, 3.0, 2000)
    assert wc == gc and len(want) == len(got)
    assert np.array_equal(want.view(np.uint8), got.view(np.uint8))


def test_harris_norm_factors(orc):
    for g in (3, 5, 7):
        for b in (3, 5, 7):
            f = np.float32(255.0) * np.float32(1 << (g - 1)) * np.float32(b)
            assert np.float32(orc.lib.oracle_harris_norm_gb(g, b)) == ((f * f) * f) * f


@pytest.mark.parametrize("g", [5, 7])
def test_harris_sobel_closed_forms(orc, g):
    """an x-ramp p = x: gx = (sum of smoothing taps)^2 * (sum of k * derivative tap), gy = 0 in the
    interior; at w % 4 == 0 the packed writer rows coincide with the plane's rows"""
    w, h = 48, 20
    R = g // 2
    ramp = np.tile(np.arange(w, dtype=np.uint8), (h, 1))
    hg3, _ = orc.harris_planes(ramp, g, 3, 0.04, 0.0)
    px = hg3.reshape(h, w, 3)
    ssum = {5: 16, 7: 64}[g]
    dsum = {5: 8, 7: 32}[g]
    gx = float(ssum * dsum)
    inner = px[R:h - R, R:w - R]
    assert (inner[..., 0] == np.float32(gx * gx)).all() and (inner[..., 1] == 0).all() and (inner[..., 2] == 0).all()


def test_harris_7x7_score_shift(orc):
    """the 7x7 score stores column x's value at x - 1"""
    src = synth.scene_u8(64, 40, 5)
    _, v7 = orc.harris_planes(src, 3, 7, 0.04, -1e30)   # every window's score kept
    assert (v7[:, 0:2] == 0).all() and (v7[:, 60:] == 0).all() and (v7[:3] == 0).all() and (v7[-3:] == 0).all()
    assert (v7[3:-3, 2:60] != 0).all()
