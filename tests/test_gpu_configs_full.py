"""The configurations at the shapes the bench runs them, checked against the oracle:
* the 256-image uneven batch (Gaussian3x3 -> Sobel3x3 -> Magnitude + Phase + half-size ScaleImage) as ONE
  persistent launch, every image and every output -- with stored gradients, and with virtual gradients
  (the merged SobelMagnitudePhase rules, 4.25 B/px);
* the fused 4K vision graph with 16 frames per launch (the bench's step), frames 0, 7 and 15.
The oracle runs in a thread pool (its C functions release the GIL)."""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest
import torch

from mivisionx_amd import synth
from mivisionx_amd.device import Plane
from mivisionx_amd.pipelines import ChainBatch
from util import assert_same, sha

pytestmark = pytest.mark.gpu
WORKERS = 16


@pytest.fixture(scope="module")
def batch256(gpu):
    sizes = synth.batch_sizes()
    assert len(sizes) == 256
    planes = []
    for i, (w, h) in enumerate(sizes):
        p = Plane(w, h, "U8")
        p.t[:, :w].copy_(synth.scene_u8_device(w, h, seed=i))
        planes.append(p)
    torch.cuda.synchronize()
    hosts = [p.download() for p in planes]
    return sizes, planes, hosts


def test_256_uneven_images_one_launch_every_output(batch256, orc):
    sizes, planes, hosts = batch256
    print(f"256-image batch: {sum(w * h for w, h in sizes) / 1e6:.1f} Mpx in one launch", flush=True)
    b = ChainBatch(planes).prepare(4)
    b.run()
    torch.cuda.synchronize()

    def check(i):
        img = hosts[i]
        h, w = img.shape
        gx, gy, mag, ph = orc.chain1(img)
        want = dict(gx=gx, gy=gy, mag=mag, phase=ph, half=orc.scale(img, w // 2, h // 2))
        got = b.results(i)
        for k in want:
            assert_same(got[k], want[k], f"image {i} ({w}x{h}) {k}")
        return i

    with ThreadPoolExecutor(WORKERS) as ex:
        assert sorted(ex.map(check, range(len(sizes)))) == list(range(256))


def test_256_uneven_images_virtual_gradients(batch256, orc):
    sizes, planes, hosts = batch256
    b = ChainBatch(planes, virtual_gradients=True).prepare(4)
    b.run()
    torch.cuda.synchronize()

    def check(i):
        img = hosts[i]
        h, w = img.shape
        wm, wp = orc.sobel_merged(img, 0, blur=True)
        got = b.results(i)
        assert_same(got["mag"], wm, f"image {i} ({w}x{h}) merged magnitude")
        assert_same(got["phase"], wp, f"image {i} ({w}x{h}) merged phase")
        assert_same(got["half"], orc.scale(img, w // 2, h // 2), f"image {i} half")
        return i

    with ThreadPoolExecutor(WORKERS) as ex:
        assert len(list(ex.map(check, range(len(sizes))))) == 256


def test_fused_4k_sixteen_frames_per_launch(gpu, orc):
    """the bench's step shape: 16 resident 4K frames through mvx_VisionGraph_run (one tile-pass launch for
    all 16); frames 0, 7 and 15 against the oracle (the frames are built as the bench builds them)"""
    from mivisionx_amd.pipelines import FusedVisionGraph
    W, H = 3840, 2160
    params = dict(lower=80, upper=100, k=0.04, thr=0.0005, dist=5.0, capacity=65536)
    base = [synth.scene_rgb(W, H, s) for s in (0, 1)]
    G = FusedVisionGraph(W, H, W // 2, H // 2, 5, frames=16, **params)
    inputs = {}
    for f, fr in enumerate(G.frames):
        img = base[f % 2] if f < 2 else np.roll(base[f % 2], (37 * f, 53 * f), axis=(0, 1))
        fr.rgb.upload(img)
        if f in (0, 7, 15):
            inputs[f] = img
    print("fused 4K, 16 frames per step", flush=True)
    G.run()
    torch.cuda.synchronize()

    def check(f):
        got = G.frames[f].results()
        want = orc.graph2(inputs[f], W // 2, H // 2, 5, params["lower"], params["upper"], params["k"], params["thr"],
                          params["dist"], params["capacity"])
        tag = f"4K frame {f} of 16"
        assert_same(got["scaled"], want["scaled"], "scaled " + tag)
        for l, (a, b) in enumerate(zip(got["pyramid"], want["pyramid"])):
            assert_same(a, b, f"pyramid {l} " + tag)
        assert_same(got["edges"], want["edges"], "edges " + tag)
        assert got["count"] == want["count"], "corner count " + tag
        assert_same(got["corners"], want["corners"], "corners " + tag)
        return f

    with ThreadPoolExecutor(3) as ex:
        assert sorted(ex.map(check, sorted(inputs))) == [0, 7, 15]

    # the border unions run concurrently inside the tile pass: every later batch gives every frame the
    # same edges and corners as the first one
    first = [(sha(fr.edges.download()), int(fr.count.item())) for fr in G.frames]
    for rep in range(4):
        G.run()
        torch.cuda.synchronize()
        assert [(sha(fr.edges.download()), int(fr.count.item())) for fr in G.frames] == first, f"batch {rep + 2}"
