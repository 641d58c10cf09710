"""Pins the oracle's pyramidal Lucas-Kanade restatement (CPU only): golden hashes of the reference
engine's tracked keypoints (window sizes 5-15, 2-4 levels, ITERATIONS and BOTH termination, initial
estimates, a node run three times), the live reference engine when oracle/_ref is built, and the
closed form of the workload: a translated copy is tracked by the translation."""
import json
import os

import numpy as np
import pytest

from mivisionx_amd import synth
from pyoracle import KEYPOINT
from util import sha

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BOTH, EPS = 0xF002, 0xF001


@pytest.fixture(scope="module")
def glk():
    with open(os.path.join(ROOT, "tests", "golden", "vision_golden_lk.json")) as f:
        return json.load(f)


def _estimate(kp):
    est = kp.copy()
    est["x"] += 2
    est["y"] -= 1
    return est


def test_lk_matches_golden(orc, glk):
    for c in glk["cases"]:
        if c["w"] * c["h"] > 1_000_000:
            continue   # the 1920x1080 case is checked on the GPU side
        old, new, t = synth.lk_pair(c["w"], c["h"], c["pair_seed"])
        assert sha(old) == c["old_sha256"] and sha(new) == c["new_sha256"]
        kp = synth.lk_keypoints(c["w"], c["h"], c["points"], c["keypoint_seed"], c["margin"])
        assert sha(kp) == c["keypoints_sha256"]
        op, np_ = orc.pyramid(old, c["levels"]), orc.pyramid(new, c["levels"])
        state = orc.lk_state(c["w"], c["h"], c["points"])
        for _ in range(c["reps"]):
            got = orc.lk(op, np_, kp, _estimate(kp), c["termination"], glk["epsilon"], glk["iterations"], c["use_estimate"],
                         c["window"], state=state)
        assert sha(got) == c["result_sha256"], (c["w"], c["h"], c["levels"], c["window"], c["termination"], c["reps"])


def test_lk_matches_live_reference(orc, ref):
    w, h, levels = 400, 300, 3
    old, new, _ = synth.lk_pair(w, h, 9)
    kp = synth.lk_keypoints(w, h, 700, 10, margin=6)
    kp["tracking_status"][::17] = 0          # points already lost are copied through
    op, np_ = orc.pyramid(old, levels), orc.pyramid(new, levels)
    for win in (3, 8, 11):
        for term in (0xF000, BOTH):
            want = ref.lk(old, new, levels, kp, _estimate(kp), term, 0.02, 12, 1, win)
            got = orc.lk(op, np_, kp, _estimate(kp), term, 0.02, 12, 1, win)
            assert (got.view(np.uint8) == want.view(np.uint8)).all(), (win, term)


def test_lk_recovers_the_translation(orc):
    """the workload's own closed form: tracked points move by the seeded translation (within 0.5 px)"""
    w, h = 640, 480
    old, new, (tx, ty) = synth.lk_pair(w, h, 11)
    kp = synth.lk_keypoints(w, h, 1500, 12)
    got = orc.lk(orc.pyramid(old, 4), orc.pyramid(new, 4), kp, kp, BOTH, 0.01, 20, 0, 9)
    ok = got["tracking_status"] == 1
    assert ok.mean() > 0.95
    dx, dy = np.median(got["x"][ok] - kp["x"][ok]), np.median(got["y"][ok] - kp["y"][ok])
    assert abs(dx - tx) <= 0.5 + 1e-9 and abs(dy - ty) <= 0.5 + 1e-9, ((dx, dy), (tx, ty))


def test_lk_epsilon_only_terminates(orc):
    """EPSILON-only termination: the reference loops without bound; the restatement caps it"""
    w, h = 320, 240
    old, new, _ = synth.lk_pair(w, h, 13)
    kp = synth.lk_keypoints(w, h, 300, 14)
    got = orc.lk(orc.pyramid(old, 3), orc.pyramid(new, 3), kp, kp, EPS, 0.01, 20, 0, 9)
    assert got.dtype == KEYPOINT and (got["tracking_status"] == 1).mean() > 0.9
