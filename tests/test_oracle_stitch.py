"""Closed forms for the stitch merge restatement (Loom's merge has no CPU form in the reference, so
these pin it): one camera at weight 255 copies its image; complementary weights over equal images give
the image back; zero weights give black; weights scale linearly."""
import numpy as np

from mivisionx_amd import synth


def _rgb(w, h, seed):
    return np.stack([synth.scene_u8(w, h, seed + c) for c in range(3)], -1)


def test_stitch_merge_closed_forms(orc):
    w, h = 64, 40
    a, b = _rgb(w, h, 1), _rgb(w, h, 9)
    full, zero = np.full((h, w), 255, np.uint8), np.zeros((h, w), np.uint8)
    assert np.array_equal(orc.stitch_merge([a], [full]), a)
    assert np.array_equal(orc.stitch_merge([a, b], [full, zero]), a)
    half1, half2 = np.full((h, w), 128, np.uint8), np.full((h, w), 127, np.uint8)
    assert np.array_equal(orc.stitch_merge([a, a], [half1, half2]), a)
    assert not orc.stitch_merge([a, b], [zero, zero]).any()
    mixed = orc.stitch_merge([a, b], [half1, half2]).astype(int)
    want = (a.astype(float) * 128 + b.astype(float) * 127) / 255
    assert np.abs(mixed - want).max() <= 0.5 + 1e-4


# ------------------------------------------------------------------ Loom's camera-selection merge ----

def _sel(h, groups, code, ids1=(31, 31, 31), ids2=(31, 31, 31)):
    id0 = np.full((h, groups), code, np.uint8)
    id1 = np.full((h, groups), ids1[0] | (ids1[1] << 5) | (ids1[2] << 10), np.uint16)
    id2 = np.full((h, groups), ids2[0] | (ids2[1] << 5) | (ids2[2] << 10), np.uint16)
    return id0, id1, id2


def test_loom_merge_selection_closed_forms(orc):
    """single camera = copy (weights ignored); 31 = untouched; complementary weights over equal images copy;
    an id beyond the cameras or 31 in a list contributes nothing"""
    w, h, n = 64, 6, 3
    cams = [_rgb(w, h, 20 + 3 * k) for k in range(n)]
    stacked = np.concatenate(cams, 0)
    rng = np.random.default_rng(0)
    wts = rng.integers(0, 256, (n * h, w), dtype=np.uint8)
    g = w // 8
    for k in range(n):
        assert np.array_equal(orc.stitch_loom(stacked, wts, *_sel(h, g, k)), cams[k])
    pre = np.full((h, w, 3), 77, np.uint8)
    assert np.array_equal(orc.stitch_loom(stacked, wts, *_sel(h, g, 31), dst=pre.copy()), pre)
    assert np.array_equal(orc.stitch_loom(stacked, wts, *_sel(h, g, 5), dst=pre.copy()), pre)   # no camera 5
    same = np.concatenate([cams[0]] * n, 0)
    a = rng.integers(0, 256, (h, w), dtype=np.uint8)
    comp = np.concatenate([a, (255 - a).astype(np.uint8), a], 0)
    assert np.array_equal(orc.stitch_loom(same, comp, *_sel(h, g, 128, (0, 1, 2))), cams[0])   # the third is not read at 128
    full = np.full((n * h, w), 255, np.uint8)
    assert np.array_equal(orc.stitch_loom(stacked, full, *_sel(h, g, 128, (1, 31, 0))), cams[1])


def test_loom_merge_blend_count_thresholds(orc):
    """camId0 > 128 adds camId1's third camera, > 129 / > 130 / > 131 camId2's first / second / third;
    each camera contributes weight * 0.003922 * pixel, summed in list order in float32"""
    w, h, n = 32, 2, 6
    vals = [3, 10, 30, 60, 100, 40]
    stacked = np.concatenate([np.full((h, w, 3), v, np.uint8) for v in vals], 0)
    wts = np.full((n * h, w), 200, np.uint8)
    f = np.float32(0.003922)
    for code, count in ((40, 2), (128, 2), (129, 3), (130, 4), (131, 5), (132, 6), (200, 6)):
        out = orc.stitch_loom(stacked, wts, *_sel(h, w // 8, code, (0, 1, 2), (3, 4, 5)))
        acc = np.float32(0)
        for v in vals[:count]:
            acc = np.float32(acc + np.float32(np.float32(200) * f) * np.float32(v))
        assert (out == int(np.rint(acc))).all(), (code, count, out[0, 0], acc)


def test_loom_fused_equals_remap_then_merge(orc):
    """the one-pass form (each camera sampled through its 13.3 table) equals the pinned per-channel remap
    (constant border 0) of every camera followed by the merge, on a small rig"""
    from mivisionx_amd import synth
    rig = synth.stitch_rig(W=384, H=192, sw=96, sh=54, device="cpu")
    W, H, n = rig["W"], rig["H"], len(rig["images"])
    fixed = np.concatenate([orc.remap_fixed(m.numpy(), rig["sw"], rig["sh"]) for m in rig["maps"]], 0)
    stacked = np.concatenate([np.stack([orc.remap(np.ascontiguousarray(img[..., c]), m.numpy(), True, 0) for c in range(3)], -1)
                              for img, m in zip(rig["images"], rig["maps"])], 0)
    args = (rig["weights"].numpy(), rig["camId0"].numpy(), rig["camId1"].numpy().view(np.uint16), rig["camId2"].numpy().view(np.uint16))
    two = orc.stitch_loom(stacked, *args)
    one = orc.stitch_loom_remap(rig["images"], fixed.reshape(n * H, W, 2), *args, workers=4)
    assert np.array_equal(one, two)
    assert (rig["camId0"].numpy() != 31).any() and (rig["camId0"].numpy() > 31).any()
