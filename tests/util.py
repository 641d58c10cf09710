import hashlib

import numpy as np


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def assert_same(got, want, what=""):
    """bit-exact comparison with a readable report of the first mismatches"""
    got = np.asarray(got)
    want = np.asarray(want)
    assert got.shape == want.shape, f"{what}: shape {got.shape} != {want.shape}"
    bad = np.argwhere(got != want)
    if bad.size:
        pts = [(tuple(int(v) for v in p), got[tuple(p)].item(), want[tuple(p)].item()) for p in bad[:8]]
        raise AssertionError(f"{what}: {len(bad)} of {got.size} differ; first (index, got, want): {pts}")
