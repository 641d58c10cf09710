"""CPU checks for the uneven-size batch: the size generator, the device scene generator against
the numpy one, the oracle's chain and half-scale against the reference engine's hashes, and host-side
validation of mvx_ChainBatch_prepare (no GPU needed)."""
import ctypes
import json
import os

import numpy as np
import pytest

from mivisionx_amd import synth
from util import sha

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_batch_sizes():
    s = synth.batch_sizes()
    assert len(s) == 256
    assert all(w % 16 == 0 and h % 16 == 0 and 320 <= w <= 4096 and 240 <= h <= 2160 for w, h in s)
    assert s == synth.batch_sizes()
    assert len(set(s)) > 200
    # the spread covers both ends of both ranges
    assert min(w for w, _ in s) < 480 and max(w for w, _ in s) > 3900
    assert min(h for _, h in s) < 360 and max(h for _, h in s) > 2000


@pytest.mark.parametrize("w,h,seed", [(336, 240, 0), (1000, 512, 7)])
def test_device_scene_generator_matches_numpy_on_cpu(w, h, seed):
    a = synth.scene_u8(w, h, seed)
    b = synth.scene_u8_device(w, h, seed, device="cpu").numpy()
    d = np.abs(a.astype(np.int32) - b)
    assert d.max() <= 1 and (d != 0).mean() < 1e-4


def test_oracle_chain_and_half_vs_reference_hashes(orc):
    with open(os.path.join(ROOT, "tests", "golden", "vision_golden_chain_batch.json")) as f:
        c = json.load(f)["cases"][1]
    img = synth.scene_u8(c["width"], c["height"], seed=c["seed"])
    assert sha(img) == c["input_sha256"]
    gx, gy, mag, ph = orc.chain1(img)
    half = orc.scale(img, c["width"] // 2, c["height"] // 2)
    for k, v in dict(gx=gx, gy=gy, mag=mag, phase=ph, half=half).items():
        assert sha(v) == c[k + "_sha256"], k


class _Image(ctypes.Structure):
    _fields_ = [("src", ctypes.c_void_p), ("width", ctypes.c_uint32), ("height", ctypes.c_uint32),
                ("srcStride", ctypes.c_uint32), ("gx", ctypes.c_void_p), ("gxStride", ctypes.c_uint32),
                ("gy", ctypes.c_void_p), ("gyStride", ctypes.c_uint32), ("mag", ctypes.c_void_p),
                ("magStride", ctypes.c_uint32), ("phase", ctypes.c_void_p), ("phaseStride", ctypes.c_uint32),
                ("half", ctypes.c_void_p), ("halfStride", ctypes.c_uint32)]


@pytest.mark.parametrize("bad", ["odd_width", "odd_height", "stride", "align", "count", "tile"])
def test_prepare_validates_on_host(bad):
    import mivisionx_amd
    from mivisionx_amd.pipelines import ChainPlan
    lib = mivisionx_amd.kernels()
    p = 1 << 20   # never dereferenced: validation fails first
    w, h = 64, 32
    im = _Image(p, w, h, 64, p, 128, p, 128, p, 128, p, 64, p, 32)
    count, tile = 1, 0
    if bad == "odd_width":
        im.width = 63
    elif bad == "odd_height":
        im.height = 31
    elif bad == "stride":
        im.gxStride = 120
    elif bad == "align":
        im.mag = p + 8
    elif bad == "count":
        count = 0
    else:
        tile = 5
    plan = ChainPlan()
    rc = lib.mvx_ChainBatch_prepare(ctypes.byref(im), count, tile, 0, ctypes.c_void_p(p), ctypes.c_size_t(4096),
                                    ctypes.byref(plan))
    assert rc == -10
