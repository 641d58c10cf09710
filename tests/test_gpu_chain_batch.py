"""GPU parity of the batched fused chain (Gaussian3x3 -> Sobel GXY -> Magnitude + Phase, ScaleImage
to half size) over images of uneven sizes in one persistent launch: bit-exact against the oracle's
node-by-node chain, against the per-node HIP kernels, against hashes of the reference CPU engine
(tests/golden/vision_golden_chain_batch.json), for every tile shape, through a captured HIP graph
replayed several times, and after the autotuner's pick."""
import json
import os

import numpy as np
import pytest
import torch

from mivisionx_amd import synth
from mivisionx_amd.device import Plane
from mivisionx_amd.pipelines import CHAIN_TILES, ChainBatch
from util import assert_same, sha

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# uneven sizes: widths not multiples of 8 or 16, heights down to 2, one image narrower than a tile
SMALL = [(34, 18), (50, 2), (322, 242), (16, 4), (130, 66), (64, 32), (2, 6), (98, 100), (640, 480), (18, 130)]


def _oracle(orc, img):
    h, w = img.shape
    gx, gy, mag, ph = orc.chain1(img)
    return dict(gx=gx, gy=gy, mag=mag, phase=ph, half=orc.scale(img, w // 2, h // 2))


@pytest.fixture(scope="module")
def small_batch(gpu, orc):
    imgs = [synth.scene_u8(w, h, seed=50 + i) if i % 2 else synth.uniform_u8(w, h, seed=50 + i)
            for i, (w, h) in enumerate(SMALL)]
    want = [_oracle(orc, a) for a in imgs]
    return imgs, want


def _check(batch, want, what):
    torch.cuda.synchronize()
    for i, wv in enumerate(want):
        got = batch.results(i)
        for k in wv:
            assert_same(got[k], wv[k], f"{what} image {i} {k}")


@pytest.mark.parametrize("tile", range(len(CHAIN_TILES)))
def test_small_uneven_every_tile_vs_oracle(small_batch, tile):
    imgs, want = small_batch
    print(f"chain batch: {len(imgs)} small images, tile {CHAIN_TILES[tile]}", flush=True)
    b = ChainBatch([Plane.from_array(a) for a in imgs]).prepare(tile)
    b.run()
    _check(b, want, f"tile {CHAIN_TILES[tile]}")


@pytest.mark.parametrize("per_cu", [1, 2, 0])
def test_grid_sizes(small_batch, per_cu):
    imgs, want = small_batch
    b = ChainBatch([Plane.from_array(a) for a in imgs]).prepare(4, per_cu)
    b.run()
    _check(b, want, f"{per_cu} workgroups per CU")


def test_per_image_launches_match(small_batch):
    imgs, want = small_batch
    b = ChainBatch([Plane.from_array(a) for a in imgs])
    b.run_per_image()
    _check(b, want, "per-image launches")


def test_graph_replays(small_batch):
    """replays of the captured launch, with eager work and synchronisation between them"""
    imgs, want = small_batch
    b = ChainBatch([Plane.from_array(a) for a in imgs]).prepare(2)
    b.capture()
    for rep in range(3):
        b.clear_outputs()
        torch.cuda.synchronize()
        b.launch_graph()
        _check(b, want, f"graph replay {rep}")
        torch.cuda.synchronize()


def test_batch_workload_images_vs_reference_hashes(gpu):
    """the first images of the 256-image uneven batch (the uneven-size benchmark workload), with
    outputs hashed by the reference CPU engine"""
    with open(os.path.join(ROOT, "tests", "golden", "vision_golden_chain_batch.json")) as f:
        golden = json.load(f)
    cases = golden["cases"]
    imgs = [synth.scene_u8(c["width"], c["height"], seed=c["seed"]) for c in cases]
    for c, a in zip(cases, imgs):
        assert sha(a) == c["input_sha256"], "generator drifted"
    b = ChainBatch([Plane.from_array(a) for a in imgs]).prepare(4)
    b.run()
    torch.cuda.synchronize()
    for i, c in enumerate(cases):
        got = b.results(i)
        for k in ("gx", "gy", "mag", "phase", "half"):
            assert sha(got[k]) == c[k + "_sha256"], f"image {i} ({c['width']}x{c['height']}) {k}"


def test_autotune_then_batch_vs_per_node(gpu):
    """a slice of the uneven batch: the autotuned launch against the per-node kernels"""
    sizes = synth.batch_sizes()[:24]
    srcs = [Plane(w, h, "U8") for (w, h) in sizes]
    for i, (p, (w, h)) in enumerate(zip(srcs, sizes)):
        p.t[:, :w].copy_(synth.scene_u8_device(w, h, seed=i))
    b = ChainBatch(srcs)
    table = b.autotune(reps=2)
    assert all(v > 0 for v in table.values()), table
    print(f"autotune pick: tile {b.plan.tileWidth}x{b.plan.tileHeight}, {b.plan.workgroups} workgroups", flush=True)
    torch.cuda.synchronize()
    got = [b.results(i) for i in range(len(srcs))]
    b.clear_outputs()
    b.run_per_image()
    torch.cuda.synchronize()
    for i in range(len(srcs)):
        want = b.results(i)
        for k in want:
            assert_same(got[i][k], want[k], f"image {i} {sizes[i]} {k}")
