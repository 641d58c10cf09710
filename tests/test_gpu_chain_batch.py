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


def test_batchpd_layout(gpu, orc):
    """the batchPD front end (images stacked in tall planes, per-image widths/heights arrays) gives the
    same per-image results as the oracle's node chain.  Neighbour reads are linear in the tall plane's
    stride, so the widths here are not multiples of 16: then the bytes left of column 0 and right of
    column w-1 are zero padding in both layouts"""
    import ctypes
    from mivisionx_amd import kernels
    from mivisionx_amd.pipelines import ChainPlan
    dims = [(34, 18), (322, 242), (18, 4), (130, 66), (2, 6), (98, 100)]
    n = len(dims)
    maxW, maxH = max(w for w, _ in dims), max(h for _, h in dims)
    al = lambda b: (b + 15) & ~15
    src = torch.zeros((n * maxH, al(maxW)), dtype=torch.uint8, device="cuda")
    imgs = []
    for i, (w, h) in enumerate(dims):
        a = synth.scene_u8(w, h, seed=70 + i) if w >= 16 and h >= 16 else synth.uniform_u8(w, h, seed=70 + i)
        imgs.append(a)
        src[i * maxH:i * maxH + h, :w] = torch.from_numpy(a).cuda()
    gx = torch.zeros((n * maxH, al(2 * maxW)), dtype=torch.uint8, device="cuda")
    gy, mag = torch.zeros_like(gx), torch.zeros_like(gx)
    ph = torch.zeros((n * maxH, al(maxW)), dtype=torch.uint8, device="cuda")
    half = torch.zeros((n * (maxH // 2), al(maxW // 2)), dtype=torch.uint8, device="cuda")
    K = kernels()
    K.mvx_ChainBatch_workspace_size.restype = ctypes.c_size_t
    wsz = K.mvx_ChainBatch_workspace_size(n)
    ws = torch.zeros(int(wsz) + 256, dtype=torch.uint8, device="cuda")
    U32 = ctypes.c_uint32 * n
    plan = ChainPlan()
    p = lambda t: ctypes.c_void_p(t.data_ptr())
    rc = K.mvx_ChainBatch_prepare_batchPD(p(src), src.shape[1], maxH, U32(*[w for w, _ in dims]), U32(*[h for _, h in dims]), n,
                                          p(gx), gx.shape[1], p(gy), gy.shape[1], p(mag), mag.shape[1], p(ph), ph.shape[1],
                                          p(half), half.shape[1], maxH // 2, 4, 0, p(ws), ctypes.c_size_t(int(wsz)),
                                          ctypes.byref(plan))
    assert rc == 0
    assert K.mvx_ChainBatch_run(ctypes.c_void_p(torch.cuda.current_stream().cuda_stream), ctypes.byref(plan)) == 0
    torch.cuda.synchronize()
    G, Y, M, P_, H = (t.cpu().numpy() for t in (gx, gy, mag, ph, half))
    for i, (w, h) in enumerate(dims):
        want = _oracle(orc, imgs[i])
        r0 = i * maxH
        assert_same(G[r0:r0 + h].view(np.int16)[:, :w], want["gx"], f"batchPD gx {i}")
        assert_same(Y[r0:r0 + h].view(np.int16)[:, :w], want["gy"], f"batchPD gy {i}")
        assert_same(M[r0:r0 + h].view(np.int16)[:, :w], want["mag"], f"batchPD mag {i}")
        assert_same(P_[r0:r0 + h, :w], want["phase"], f"batchPD phase {i}")
        hr = i * (maxH // 2)
        assert_same(H[hr:hr + h // 2, :w // 2], want["half"], f"batchPD half {i}")
