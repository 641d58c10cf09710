"""mvx_NonMaxSupp_XY_ANY_3x3_NoAlloc against mvx_NonMaxSupp_XY_ANY_3x3 (the workspace form) through
the C ABI: two calls in a row on one GPU, the second on a smaller image, so the second reads the
library's scratch array as the first left it; bit-exact item lists and counts.  Also the size limit:
an image needing more scratch than the array holds is refused before anything is launched."""
import ctypes

import numpy as np
import pytest

from util import assert_same

pytestmark = pytest.mark.gpu


def _score_plane(w, h, seed):
    """a float plane with many equal neighbours (small integers, zero plateaus), so both the >= and
    the > sides of the suppression rule decide"""
    from mivisionx_amd.device import Plane
    rng = np.random.default_rng(seed)
    v = rng.integers(0, 6, size=(h, w)).astype(np.float32)
    v[rng.random((h, w)) < 0.3] = 0.0
    return Plane(w, h, "F32").upload(v)


def _run(K, fn, w, h, plane, capacity, with_ws):
    import torch
    from mivisionx_amd.device import stream_handle
    items = torch.full((8 * capacity + 64,), 0xA5, dtype=torch.uint8, device="cuda")
    count = torch.full((1,), 0x7FFFFFFF, dtype=torch.int32, device="cuda")
    args = [stream_handle(), ctypes.c_uint32(capacity), ctypes.c_void_p(items.data_ptr()), ctypes.c_void_p(count.data_ptr()),
            ctypes.c_uint32(w), ctypes.c_uint32(h), plane.ptr, ctypes.c_uint32(plane.stride)]
    ws = None
    if with_ws:
        size = K.mvx_NonMaxSupp_XY_ANY_3x3_workspace_size(w, h)
        ws = torch.full((size,), 0x5A, dtype=torch.uint8, device="cuda")
        args += [ctypes.c_void_p(ws.data_ptr()), ctypes.c_size_t(size)]
    rc = fn(*args)
    assert rc == 0, f"{'workspace' if with_ws else 'NoAlloc'} form returned {rc} at {w}x{h}"
    torch.cuda.synchronize()
    n = int(count.item())
    return n, items[: 8 * min(n, capacity)].cpu().numpy()


def test_noalloc_matches_workspace_form_two_calls(gpu):
    from mivisionx_amd import kernels
    K = kernels()
    K.mvx_NonMaxSupp_XY_ANY_3x3_workspace_size.restype = ctypes.c_size_t
    for w, h, seed, capacity in ((1920, 1080, 1, 1 << 20), (640, 480, 2, 1 << 18)):
        print(f"NoAlloc NMS {w}x{h} seed {seed}", flush=True)
        plane = _score_plane(w, h, seed)
        n_ref, ref = _run(K, K.mvx_NonMaxSupp_XY_ANY_3x3, w, h, plane, capacity, True)
        n_got, got = _run(K, K.mvx_NonMaxSupp_XY_ANY_3x3_NoAlloc, w, h, plane, capacity, False)
        assert n_ref > 1000, f"{w}x{h}: only {n_ref} maxima; the input does not exercise the compaction"
        assert n_got == n_ref, f"{w}x{h}: count {n_got} != {n_ref}"
        assert_same(got, ref, f"NoAlloc NMS items {w}x{h}")


def test_noalloc_truncates_at_capacity_like_workspace_form(gpu):
    from mivisionx_amd import kernels
    K = kernels()
    K.mvx_NonMaxSupp_XY_ANY_3x3_workspace_size.restype = ctypes.c_size_t
    plane = _score_plane(800, 600, 3)
    n_ref, ref = _run(K, K.mvx_NonMaxSupp_XY_ANY_3x3, 800, 600, plane, 500, True)
    n_got, got = _run(K, K.mvx_NonMaxSupp_XY_ANY_3x3_NoAlloc, 800, 600, plane, 500, False)
    assert n_ref == n_got == 500
    assert_same(got, ref, "NoAlloc NMS items truncated at capacity 500")


def test_noalloc_refuses_images_beyond_its_scratch(gpu):
    import torch
    from mivisionx_amd import kernels
    from mivisionx_amd.device import Plane, stream_handle
    K = kernels()
    K.mvx_NonMaxSupp_XY_ANY_3x3_workspace_size.restype = ctypes.c_size_t
    w, h = 16384, 4096
    assert K.mvx_NonMaxSupp_XY_ANY_3x3_workspace_size(w, h) > (1 << 20)
    plane = Plane(w, h, "F32")
    count = torch.full((1,), 12345, dtype=torch.int32, device="cuda")
    items = torch.zeros(64, dtype=torch.uint8, device="cuda")
    rc = K.mvx_NonMaxSupp_XY_ANY_3x3_NoAlloc(stream_handle(), ctypes.c_uint32(8), ctypes.c_void_p(items.data_ptr()),
                                             ctypes.c_void_p(count.data_ptr()), ctypes.c_uint32(w), ctypes.c_uint32(h),
                                             plane.ptr, ctypes.c_uint32(plane.stride))
    torch.cuda.synchronize()
    assert rc == -8, f"expected MVX_ERROR_NO_MEMORY (-8), got {rc}"
    assert int(count.item()) == 12345, "a refused call must launch nothing"
