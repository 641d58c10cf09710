"""Write tests/golden/vision_golden_warp.json from the reference's own CPU engine (oracle/_ref):
WarpAffine / WarpPerspective (bilinear, UNDEFINED and CONSTANT borders) and Remap (bilinear, both
borders) on seeded scenes with seeded near-identity matrices and barrel maps.

Pixels whose reference value depends on memory past the source allocation (the oracle flags them as
unpinned) are zeroed in both the reference output and the oracle output before hashing; their count
is recorded.  Inputs are regenerated from their seeds by mivisionx_amd.synth.

    python tests/golden/make_golden_warp.py
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
from pyoracle import Oracle, Reference  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "vision_golden_warp.json")
SIZES = [(333, 77), (640, 480), (64, 48), (37, 29), (1920, 1080)]


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main():
    ref, orc = Reference(), Oracle()
    cases = []
    for (w, h) in SIZES:
        src = synth.scene_u8(w, h, 21)
        for seed in (1, 2):
            aff, per = synth.warp_matrices(w, h, seed)
            dw, dh = (w, h) if seed == 1 else (w + 3, max(3, h - 5))
            for border in (None, 0, 77):
                c = {"w": w, "h": h, "seed": seed, "dw": dw, "dh": dh, "border": border, "src_sha256": sha(src),
                     "affine": aff.tolist(), "perspective": per.tolist()}
                got, unp = orc.warp_affine(src, dw, dh, aff, border is not None, border or 0)
                want = ref.warp(0, src, dw, dh, aff, border is not None, border or 0)
                want[unp] = 0
                c["affine_sha256"], c["affine_unpinned"] = sha(want), int(unp.sum())
                got, unp = orc.warp_perspective(src, dw, dh, per, border or 0)
                want = ref.warp(1, src, dw, dh, per, border is not None, border or 0)
                want[unp] = 0
                c["perspective_sha256"], c["perspective_unpinned"] = sha(want), int(unp.sum())
                cases.append(c)
        if w * h < 100000:
            # exact integer / half-pixel shifts: the perspective floor quirk (trunc - (x < 0) maps -1.0 to -2)
            # and the mask edges
            for name, (tx, ty) in (("shift_-1_-2", (-1.0, -2.0)), ("shift_0.5", (0.5, 0.5)), ("identity", (0.0, 0.0))):
                aff = np.array([1, 0, 0, 1, tx, ty], np.float32)
                per = np.array([1, 0, 0, 0, 1, 0, tx, ty, 1], np.float32)
                for border in (None, 9):
                    c = {"w": w, "h": h, "seed": name, "dw": w, "dh": h, "border": border, "src_sha256": sha(src),
                         "affine": aff.tolist(), "perspective": per.tolist()}
                    for kind, m in ((0, aff), (1, per)):
                        key = "affine" if kind == 0 else "perspective"
                        if kind == 0:
                            got, unp = orc.warp_affine(src, w, h, m, border is not None, border or 0)
                        else:
                            got, unp = orc.warp_perspective(src, w, h, m, border or 0)
                        want = ref.warp(kind, src, w, h, m, border is not None, border or 0)
                        want[unp] = 0
                        c[key + "_sha256"], c[key + "_unpinned"] = sha(want), int(unp.sum())
                    cases.append(c)
        for k1 in (-0.2, 0.15):
            mp = synth.barrel_map(w, h, k1)
            for border in (None, 200):
                want = ref.remap(src, mp, border is not None, border or 0)
                cases.append({"w": w, "h": h, "k1": k1, "border": border, "src_sha256": sha(src), "map_sha256": sha(mp),
                              "remap_sha256": sha(want)})
    with open(OUT, "w") as f:
        json.dump({"generator": "tests/golden/make_golden_warp.py", "reference": "oracle/_ref (reference AGO CPU engine)",
                   "scene_seed": 21, "cases": cases}, f, indent=1)
    print(f"wrote {OUT}: {len(cases)} cases")


if __name__ == "__main__":
    main()
