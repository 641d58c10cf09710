"""The merged Sobel -> Magnitude / Phase kernels of the oracle (HafCpu_SobelMagnitudePhase_S16U8_U8_3x3,
HafCpu_SobelMagnitude_S16_U8_3x3, HafCpu_SobelPhase_U8_U8_3x3 as chosen by ago_drama_remove.cpp:155-182),
pinned three ways: golden hashes written by the reference engine, live comparison with oracle/_ref, and
closed forms computed with exact integer square roots and from the unmerged Sobel gradients."""
import json
import math
import os

import numpy as np
import pytest

from mivisionx_amd import synth
from util import assert_same, sha

HERE = os.path.dirname(os.path.abspath(__file__))


def _golden():
    with open(os.path.join(HERE, "golden", "vision_golden_sobel_merged.json")) as f:
        return json.load(f)["cases"]


def _src(c):
    return synth.scene_u8(c["w"], c["h"], c["seed"]) if c["input"] == "scene" else synth.uniform_u8(c["w"], c["h"], c["seed"])


@pytest.mark.parametrize("case", _golden(), ids=lambda c: f"{c['w']}x{c['h']}-which{c['which']}-blur{c['blur']}")
def test_oracle_matches_reference_golden(orc, case):
    src = _src(case)
    assert sha(src) == case["input_sha256"], "input generator changed"
    mag, ph = orc.sobel_merged(src, case["which"], bool(case["blur"]))
    if "mag_sha256" in case:
        assert sha(mag) == case["mag_sha256"], f"magnitude differs from the reference engine (row 1 tail {mag[1, -20:].tolist()} vs {case['mag_row1']})"
    if "phase_sha256" in case:
        assert sha(ph) == case["phase_sha256"], f"phase differs from the reference engine (row 1 tail {ph[1, -20:].tolist()} vs {case['phase_row1']})"


@pytest.mark.parametrize("w,h", [(1928, 40), (333, 50), (17, 9), (64, 6), (7, 5), (1000, 3)])
@pytest.mark.parametrize("which", [0, 1, 2])
def test_oracle_matches_reference_live(orc, ref, w, h, which):
    src = synth.uniform_u8(w, h, 900 + w + h)
    for blur in (False, True):
        om, op = orc.sobel_merged(src, which, blur)
        rm, rp = ref.sobel_merged(src, which, blur)
        if which != 2:
            assert_same(om, rm, f"magnitude {w}x{h} blur={blur}")
        if which != 1:
            assert_same(op, rp, f"phase {w}x{h} blur={blur}")


def _round_half_even_sqrt(s):
    k = math.isqrt(s)
    return k + 1 if s - k * k > k else k    # sqrt(s) is never exactly k + 1/2 for an integer s


def test_merged_magnitude_pixel_rules_closed_form(orc):
    # body columns: round half to even of sqrt; postfix columns: truncated sqrt, over a grid of gradients
    for gx in range(-1020, 1021, 17):
        for gy in range(-1020, 1021, 23):
            s = gx * gx + gy * gy
            assert orc.merged_magnitude_px(gx, gy, False) == _round_half_even_sqrt(s), (gx, gy)
            assert orc.merged_magnitude_px(gx, gy, True) == math.isqrt(s), (gx, gy)


@pytest.mark.parametrize("w,h", [(1928, 12), (333, 9), (47, 5)])
def test_sobel_magnitude_phase_closed_form(orc, w, h):
    """merged magnitude = round(sqrt) on columns [0, w & ~15), trunc(sqrt) after; phase = the unmerged
    phase; rows 0 and h-1 untouched (zero)"""
    src = synth.uniform_u8(w, h, 31)
    gx, gy = orc.sobel(src)
    s = gx.astype(np.int64) ** 2 + gy.astype(np.int64) ** 2
    want = np.zeros((h, w), np.int16)
    body = w & ~15
    for y in range(1, h - 1):
        for x in range(w):
            want[y, x] = _round_half_even_sqrt(int(s[y, x])) if x < body else math.isqrt(int(s[y, x]))
    mag, ph = orc.sobel_merged(src, 0)
    assert_same(mag, want, "merged magnitude")
    unmerged_phase = orc.phase(gx, gy)
    unmerged_phase[0] = 0
    unmerged_phase[-1] = 0
    assert_same(ph, unmerged_phase, "merged phase")
    mag1, _ = orc.sobel_merged(src, 1)
    assert_same(mag1, want, "SobelMagnitude")


@pytest.mark.parametrize("w,h", [(1928, 12), (333, 9), (17, 6), (8, 5)])
def test_sobel_phase_overlapping_scratch_closed_form(orc, w, h):
    """SobelPhase writes Gx/Gy rows ALIGN16(w) bytes apart (half a row of int16), so for x >= ALIGN16(w)/2 on
    every row but the last the phase is the unmerged phase of pixel (x - ALIGN16(w)/2, y + 1)"""
    src = synth.uniform_u8(w, h, 41)
    gx, gy = orc.sobel(src)
    up = orc.phase(gx, gy)
    half = ((w + 15) & ~15) // 2
    want = np.zeros((h, w), np.uint8)
    for y in range(1, h - 1):
        for x in range(w):
            want[y, x] = up[y + 1, x - half] if (x >= half and y + 1 <= h - 2) else up[y, x]
    _, ph = orc.sobel_merged(src, 2)
    assert_same(ph, want, "SobelPhase")
    if w > 8:
        assert (want != np.where(np.arange(h)[:, None] % (h - 1) == 0, 0, up)).any(), "the quirk must be visible at this size"
