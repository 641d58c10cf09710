"""Write tests/golden/vision_golden_sobel_merged.json from the reference's own CPU engine (oracle/_ref):
graphs whose Sobel3x3 gradients are VIRTUAL and read only by Magnitude and/or Phase, which the
reference's merge rules (ago_drama_remove.cpp:155-182) turn into SobelMagnitudePhase,
SobelMagnitude or SobelPhase, with and without a virtual Gaussian3x3 in front, on seeded uniform
U8 inputs (mivisionx_amd.synth.uniform_u8) and on the scene generator at 1920x1080.

    python tests/golden/make_golden_sobel_merged.py
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

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "vision_golden_sobel_merged.json")
SIZES = [(1920, 1080, "scene"), (1928, 64, "uniform"), (333, 64, "uniform"), (17, 9, "uniform"), (1920, 31, "uniform"),
         (100, 3, "uniform"), (5, 4, "uniform"), (8, 8, "uniform")]


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def source(w, h, kind, seed):
    return synth.scene_u8(w, h, seed) if kind == "scene" else synth.uniform_u8(w, h, seed)


def main():
    ref = Reference()
    cases = []
    for i, (w, h, kind) in enumerate(SIZES):
        seed = 700 + i
        src = source(w, h, kind, seed)
        for which in (0, 1, 2):
            for blur in (0, 1):
                mag, ph = ref.sobel_merged(src, which, bool(blur))
                c = {"w": w, "h": h, "input": kind, "seed": seed, "which": which, "blur": blur, "input_sha256": sha(src)}
                if mag is not None:
                    c["mag_sha256"] = sha(mag)
                    c["mag_row1"] = [int(v) for v in mag[1, -20:]] if h > 2 else []
                if ph is not None:
                    c["phase_sha256"] = sha(ph)
                    c["phase_row1"] = [int(v) for v in ph[1, -20:]] if h > 2 else []
                cases.append(c)
    with open(OUT, "w") as f:
        json.dump({"generator": "tests/golden/make_golden_sobel_merged.py", "reference": "oracle/_ref (reference AGO CPU engine)",
                   "which": {"0": "magnitude + phase", "1": "magnitude", "2": "phase"}, "cases": cases}, f, indent=1)
    print(f"wrote {len(cases)} cases to {OUT}")


if __name__ == "__main__":
    main()
