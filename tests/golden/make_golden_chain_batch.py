"""Write tests/golden/vision_golden_chain_batch.json from the reference's own CPU engine
(oracle/_ref): for the first images of the 256-image uneven batch (mivisionx_amd.synth.batch_sizes,
image i = synth.scene_u8(w, h, seed=i)) the outputs of the Gaussian3x3 -> Sobel3x3 ->
Magnitude + Phase graph (blurred image virtual) and of ScaleImage bilinear to (w/2, h/2), as SHA-256
hashes of the planes.

    python tests/golden/make_golden_chain_batch.py
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

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "vision_golden_chain_batch.json")
COUNT = 4


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main():
    ref = Reference()
    cases = []
    for i, (w, h) in enumerate(synth.batch_sizes()[:COUNT]):
        img = synth.scene_u8(w, h, seed=i)
        gx, gy, mag, ph = ref.chain1(img)
        half = ref.scale(img, w // 2, h // 2)
        cases.append(dict(width=w, height=h, seed=i, input_sha256=sha(img), gx_sha256=sha(gx), gy_sha256=sha(gy),
                          mag_sha256=sha(mag), phase_sha256=sha(ph), half_sha256=sha(half)))
        print(w, h, flush=True)
    with open(OUT, "w") as f:
        json.dump({"generator": "synth.batch_sizes() + synth.scene_u8(w, h, seed=i)", "cases": cases}, f, indent=1)


if __name__ == "__main__":
    main()
