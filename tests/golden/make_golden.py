"""Write tests/golden/vision_golden.json from the reference's own CPU engine.

Run in the container that holds /root/reference, after `make -C oracle ref` has compiled the
reference sources into oracle/_ref.  For each case it records the SHA-256 of the seeded input and
of every output plane the reference produced, plus a few rows of each output so a failure shows
values, not just a hash.  The committed JSON is data (inputs are regenerated from their seeds by
mivisionx_amd.synth, whose hash is recorded too).

    python tests/golden/make_golden.py
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

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "vision_golden.json")


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def crop(a):
    """rows 0..2 and the last row, first and last 8 columns: catches border handling."""
    rows = sorted(set([0, 1, 2, a.shape[0] - 2, a.shape[0] - 1]) & set(range(a.shape[0])))
    return {str(r): [a[r, :8].tolist(), a[r, -8:].tolist()] for r in rows}


def rec(**planes):
    return {k: {"sha256": sha(v), "crop": crop(v)} for k, v in planes.items()}


SIZES = [(1920, 1080), (333, 77), (64, 16), (37, 5), (10, 10), (5, 3), (16, 3)]


def grad_grid(lim=1020):
    g = np.arange(-lim, lim + 1, dtype=np.int16)
    gx = np.ascontiguousarray(np.broadcast_to(g[None, :], (g.size, g.size)))
    gy = np.ascontiguousarray(np.broadcast_to(g[:, None], (g.size, g.size)))
    return gx, gy


def main():
    ref = Reference()
    cases = []
    for (w, h) in SIZES:
        for seed in (0, 7):
            src = synth.uniform_u8(w, h, seed)
            c = {"w": w, "h": h, "seed": seed, "input": "uniform_u8", "input_sha256": sha(src)}
            gx, gy, mag, ph = ref.chain1(src)
            c["chain1"] = rec(gx=gx, gy=gy, mag=mag, phase=ph)
            for which, name in enumerate(("gaussian", "box", "median")):
                c[name] = rec(out=ref.filter3x3(which, src))
            sx, sy = ref.sobel(src)
            c["sobel"] = rec(gx=sx, gy=sy)
            cases.append(c)
            print(w, h, seed, flush=True)
    gx, gy = grad_grid()
    mag, ph = ref.magnitude_phase(gx, gy)
    grid = {"lim": 1020, "magnitude": {"sha256": sha(mag)}, "phase": {"sha256": sha(ph)},
            "samples": {f"{x},{y}": [int(mag[y + 1020, x + 1020]), int(ph[y + 1020, x + 1020])]
                        for (x, y) in [(0, 0), (1, 0), (0, 1), (-1, 0), (0, -1), (1020, 1020), (-1020, 3),
                                       (5, -7), (300, 299), (-1, -1000)]}}
    doc = {"generator": "tests/golden/make_golden.py", "reference": "ROCm/MIVisionX amd_openvx CPU path (ago_haf_cpu*)",
           "cases": cases, "gradient_grid": grid}
    with open(OUT, "w") as f:
        json.dump(doc, f, indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
