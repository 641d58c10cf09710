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
