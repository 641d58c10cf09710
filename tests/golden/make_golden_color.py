"""Write tests/golden/vision_golden_color.json from the reference's own CPU engine (oracle/_ref): every
color_convert the reference runs (pyoracle.COLOR_CONVERSIONS) at two sizes with a scalar tail (a
1920x1080 frame and a 102x6 strip), plus the luma-only form of RGB / RGBX -> IYUV, NV12, YUV4 (a virtual
YUV image read only by channel_extract(Y)), on seeded uniform random planes (numpy default_rng(seed)
in the packed plane layout of pyoracle.color_planes); SHA-256 of each output plane.

    python tests/golden/make_golden_color.py
"""
import hashlib
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from pyoracle import COLOR_CONVERSIONS, Reference, color_planes  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "vision_golden_color.json")
SIZES = [(1920, 1080), (102, 6)]


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def planes(fmt, w, h, seed):
    rng = np.random.default_rng(seed)
    return [rng.integers(0, 256, shape, dtype=np.uint8) for shape in color_planes(fmt, w, h)]


def main():
    ref = Reference()
    cases = []
    for i, (src, dst) in enumerate(COLOR_CONVERSIONS):
        for j, (w, h) in enumerate(SIZES):
            seed = 9000 + 10 * i + j
            out = ref.color_convert(src, dst, w, h, planes(src, w, h, seed))
            cases.append(dict(src=src, dst=dst, w=w, h=h, seed=seed, sha=[sha(p) for p in out]))
    for i, src in enumerate(("RGB", "RGBX")):
        for k, yuv in enumerate(("IYUV", "NV12", "YUV4")):
            w, h = 102, 6
            seed = 9900 + 10 * i + k
            y = ref.color_luma(src, yuv, w, h, planes(src, w, h, seed))
            cases.append(dict(src=src, dst=yuv, w=w, h=h, seed=seed, luma_only=True, sha=[sha(y)]))
    with open(OUT, "w") as f:
        json.dump({"generator": "tests/golden/make_golden_color.py", "cases": cases}, f, indent=1)
    print(f"wrote {len(cases)} cases to {OUT}")


if __name__ == "__main__":
    main()
