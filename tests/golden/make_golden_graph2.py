"""Write tests/golden/vision_golden_graph2.json from the reference's own CPU engine.

Covers the nodes of the 4K vision graph (ColorConvert RGB->IYUV->Y, ScaleImage bilinear,
GaussianPyramid HALF, CannyEdgeDetector 3/L1, HarrisCorners 3/3) at small and ragged sizes, and the
whole graph at 3840x2160.  Inputs come from mivisionx_amd.synth (seeded; their hashes are stored).
Run where oracle/_ref has been built (`make -C oracle ref`):

    python tests/golden/make_golden_graph2.py
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

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "vision_golden_graph2.json")

GRAPH2 = dict(w=3840, h=2160, seed=0, dw=1920, dh=1080, levels=5, lower=80, upper=100, k=0.04, thr=0.0005,
              dist=5.0, capacity=65536)
NODE_SIZES = [(64, 48), (100, 38), (34, 20), (16, 10), (8, 8), (6, 6), (200, 150), (640, 480), (1920, 1080)]
HARRIS = [(0.04, 0.0005, 5.0, 5000), (0.04, 1e-6, 5.0, 50), (0.06, 1e-5, 2.0, 300), (0.04, 1e-6, 10.0, 400)]
CANNY = [(80, 100), (10, 30), (0, 255)]


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def kp_list(kp, n=6):
    return [[int(k["x"]), int(k["y"]), float(k["strength"])] for k in kp[:n]]


def main():
    ref = Reference()
    nodes = []
    for (w, h) in NODE_SIZES:
        rgb = synth.scene_rgb(w, h, 3)
        c = {"w": w, "h": h, "seed": 3, "rgb_sha256": sha(rgb)}
        y = ref.color_y(rgb)
        c["color_y"] = sha(y)
        c["scale"] = {f"{dw}x{dh}": sha(ref.scale(y, dw, dh))
                      for (dw, dh) in [((w + 1) // 2, (h + 1) // 2), (2 * w, 2 * h), (max(1, 3 * w // 4), max(1, 2 * h // 3))]}
        if min(w, h) >= 12:
            c["pyramid3"] = [sha(l) for l in ref.pyramid(y, 3)]
        c["canny"] = {f"{lo},{hi}": sha(ref.canny(y, lo, hi)) for (lo, hi) in CANNY}
        c["harris"] = []
        for (k, thr, dist, cap) in HARRIS:
            kp, n = ref.harris(y, k, thr, dist, cap)
            c["harris"].append({"k": k, "thr": thr, "dist": dist, "capacity": cap, "count": int(n),
                                "sha256": sha(kp), "head": kp_list(kp)})
        nodes.append(c)
        print("nodes", w, h, flush=True)
    g = GRAPH2
    rgb = synth.scene_rgb(g["w"], g["h"], g["seed"])
    res = ref.graph2(rgb, g["dw"], g["dh"], g["levels"], g["lower"], g["upper"], g["k"], g["thr"], g["dist"],
                     g["capacity"])
    graph = dict(g, rgb_sha256=sha(rgb), scaled=sha(res["scaled"]), pyramid=[sha(l) for l in res["pyramid"]],
                 edges=sha(res["edges"]), edge_pixels=int((res["edges"] == 255).sum()), count=int(res["count"]),
                 corners=sha(res["corners"]), corners_head=kp_list(res["corners"], 10))
    doc = {"generator": "tests/golden/make_golden_graph2.py",
           "reference": "ROCm/MIVisionX amd_openvx CPU path (ago_haf_cpu*)", "nodes": nodes, "graph2": graph}
    with open(OUT, "w") as f:
        json.dump(doc, f, indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
