"""GPU parity of the 3x3 filters, Sobel, Magnitude and Phase: HIP kernels through the C ABI versus
the oracle, bit-exact, on seeded inputs (uniform noise, the scene generator, constant and ramp
images) at ragged and full sizes, including the reference's tiny-size cases (5x3, 10x10)."""
import numpy as np
import pytest

from mivisionx_amd import synth
from mivisionx_amd.device import Plane, call, MvxError
from util import assert_same

pytestmark = pytest.mark.gpu

SIZES = [(1920, 1080), (333, 77), (64, 16), (37, 5), (10, 10), (5, 3), (16, 3), (1, 3), (4096, 17)]
NAMES = ("Gaussian_U8_U8_3x3", "Box_U8_U8_3x3", "Median_U8_U8_3x3")


def _inputs(w, h):
    yield "uniform", synth.uniform_u8(w, h, 1)
    if w * h >= 1000:
        yield "scene", synth.scene_u8(w, h, 2)
    yield "const", np.full((h, w), 200, np.uint8)


@pytest.mark.parametrize("w,h", SIZES)
def test_filters3x3(gpu, orc, w, h):
    for tag, src in _inputs(w, h):
        s = Plane.from_array(src)
        for which, name in enumerate(NAMES):
            d = Plane(w, h)
            call(name, w, h, d, d.stride, s, s.stride)
            assert_same(d.download(), orc.filter3x3(which, src), f"{name} {tag} {w}x{h}")


@pytest.mark.parametrize("w,h", SIZES)
def test_sobel(gpu, orc, w, h):
    for tag, src in _inputs(w, h):
        s = Plane.from_array(src)
        gx, gy = Plane(w, h, "S16"), Plane(w, h, "S16")
        call("Sobel_S16S16_U8_3x3_GXY", w, h, gx, gx.stride, gy, gy.stride, s, s.stride)
        ex, ey = orc.sobel(src)
        assert_same(gx.download(), ex, f"gx {tag} {w}x{h}")
        assert_same(gy.download(), ey, f"gy {tag} {w}x{h}")
        # single-output forms (the GX / GY kernels)
        gx1 = Plane(w, h, "S16")
        call("Sobel_S16S16_U8_3x3_GXY", w, h, gx1, gx1.stride, None, 0, s, s.stride)
        assert_same(gx1.download(), ex, f"gx-only {tag} {w}x{h}")


def test_magnitude_phase_exhaustive_grid(gpu, orc):
    g = np.arange(-1020, 1021, dtype=np.int16)
    gx = np.ascontiguousarray(np.broadcast_to(g[None, :], (g.size, g.size)))
    gy = np.ascontiguousarray(np.broadcast_to(g[:, None], (g.size, g.size)))
    h, w = gx.shape
    px, py = Plane.from_array(gx, "S16"), Plane.from_array(gy, "S16")
    m, p = Plane(w, h, "S16"), Plane(w, h, "U8")
    call("Magnitude_S16_S16S16", w, h, m, m.stride, px, px.stride, py, py.stride)
    call("Phase_U8_S16S16", w, h, p, p.stride, px, px.stride, py, py.stride)
    assert_same(m.download(), orc.magnitude(gx, gy), "magnitude grid")
    assert_same(p.download(), orc.phase(gx, gy), "phase grid")


@pytest.mark.parametrize("w,h", [(37, 9), (1920, 1080), (7, 3)])
def test_magnitude_phase_full_s16_range(gpu, orc, w, h):
    rng = np.random.default_rng(w * h)
    gx = rng.integers(-32768, 32768, (h, w), dtype=np.int16)
    gy = rng.integers(-32768, 32768, (h, w), dtype=np.int16)
    gx[0, :2] = [-32768, 32767]
    gy[0, :2] = [-32768, 32767]
    px, py = Plane.from_array(gx, "S16"), Plane.from_array(gy, "S16")
    m, p = Plane(w, h, "S16"), Plane(w, h, "U8")
    call("Magnitude_S16_S16S16", w, h, m, m.stride, px, px.stride, py, py.stride)
    call("Phase_U8_S16S16", w, h, p, p.stride, px, px.stride, py, py.stride)
    assert_same(m.download(), orc.magnitude(gx, gy), f"magnitude {w}x{h}")
    assert_same(p.download(), orc.phase(gx, gy), f"phase {w}x{h}")


@pytest.mark.parametrize("w,h", [(1920, 1080), (333, 77), (5, 3), (10, 10)])
def test_chain1_kernels(gpu, orc, golden, w, h):
    """Gaussian -> Sobel -> Magnitude + Phase as four launches on one stream."""
    src = synth.uniform_u8(w, h, 0)
    s, g = Plane.from_array(src), Plane(w, h)
    gx, gy, m, p = Plane(w, h, "S16"), Plane(w, h, "S16"), Plane(w, h, "S16"), Plane(w, h)
    call("Gaussian_U8_U8_3x3", w, h, g, g.stride, s, s.stride)
    call("Sobel_S16S16_U8_3x3_GXY", w, h, gx, gx.stride, gy, gy.stride, g, g.stride)
    call("Magnitude_S16_S16S16", w, h, m, m.stride, gx, gx.stride, gy, gy.stride)
    call("Phase_U8_S16S16", w, h, p, p.stride, gx, gx.stride, gy, gy.stride)
    for got, want, k in zip((gx, gy, m, p), orc.chain1(src), ("gx", "gy", "mag", "phase")):
        assert_same(got.download(), want, f"chain1.{k} {w}x{h}")


def test_bad_stride_is_rejected(gpu):
    s, d = Plane(64, 8), Plane(64, 8)
    with pytest.raises(MvxError):
        call("Gaussian_U8_U8_3x3", 64, 8, d, 40, s, s.stride)   # stride below the row width
    with pytest.raises(MvxError):
        call("Gaussian_U8_U8_3x3", 64, 8, d, 72, s, s.stride)   # not a multiple of 16
