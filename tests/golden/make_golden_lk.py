"""Write tests/golden/vision_golden_lk.json from the reference's own CPU engine (oracle/_ref):
OpticalFlowPyrLK between a seeded scene and a seeded translated copy (mivisionx_amd.synth.lk_pair),
through HALF Gaussian pyramids, for several window sizes, level counts, termination criteria and
initial-estimate settings, including a graph processed three times (the node's scratch state
carries over between runs).  EPSILON-only termination is not recorded: the reference's iteration
loop has no bound under it and did not finish on these inputs.

    python tests/golden/make_golden_lk.py
"""
import hashlib
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from mivisionx_amd import synth  # noqa: E402
from pyoracle import Reference  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "vision_golden_lk.json")
ITER, BOTH = 0xF000, 0xF002
CASES = [  # w, h, levels, points, window, termination, use_estimate, reps, keypoint margin
    (320, 240, 3, 400, 9, BOTH, 0, 1, 16),
    (640, 480, 4, 2000, 9, BOTH, 0, 1, 16),
    (640, 480, 4, 2000, 9, ITER, 0, 1, 16),
    (333, 200, 3, 300, 7, BOTH, 1, 1, 16),
    (256, 128, 2, 300, 15, BOTH, 0, 1, 4),
    (640, 480, 4, 500, 5, BOTH, 0, 1, 16),
    (640, 480, 4, 1000, 9, BOTH, 0, 3, 16),
    (1920, 1080, 4, 5000, 9, BOTH, 0, 1, 16),
]


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def estimate(kp):
    est = kp.copy()
    est["x"] += 2
    est["y"] -= 1
    return est


def main():
    ref = Reference()
    out = []
    for (w, h, levels, n, win, term, use, reps, margin) in CASES:
        old, new, t = synth.lk_pair(w, h, 3)
        kp = synth.lk_keypoints(w, h, n, 4, margin)
        got = ref.lk(old, new, levels, kp, estimate(kp), term, 0.01, 20, use, win, reps=reps)
        out.append({"w": w, "h": h, "levels": levels, "points": n, "window": win, "termination": term, "use_estimate": use,
                    "reps": reps, "margin": margin, "pair_seed": 3, "keypoint_seed": 4, "shift": list(t),
                    "old_sha256": sha(old), "new_sha256": sha(new), "keypoints_sha256": sha(kp),
                    "result_sha256": sha(got), "lost": int((got["tracking_status"] == 0).sum()),
                    "sample": [[int(r["x"]), int(r["y"]), int(r["tracking_status"])] for r in got[:8]]})
        print(w, h, levels, n, win, hex(term), use, reps, "lost", out[-1]["lost"])
    with open(OUT, "w") as f:
        json.dump({"generator": "tests/golden/make_golden_lk.py", "reference": "oracle/_ref (reference AGO CPU engine)",
                   "epsilon": 0.01, "iterations": 20, "cases": out}, f, indent=1)
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
