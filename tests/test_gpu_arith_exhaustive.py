"""Exhaustive device-side checks of the Magnitude / Phase building blocks over their whole input
domains (tests/hip/probe_arith.hip, built by build()): the reciprocal-and-FMA division used by Phase
equals IEEE single division for every integer pair 0 <= a <= b <= 32768, and the integer square-root
rounding used by Magnitude equals the reference's rint(sqrt(double)) for every |gx|, |gy| <= 32768."""
import ctypes
import os

import numpy as np

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROBE = os.path.join(ROOT, "tests", "hip", "libprobe_arith.so")


@pytest.fixture(scope="module")
def probe(gpu):
    if not os.path.exists(PROBE):
        pytest.fail(f"{PROBE} is not built (run build())")
    return ctypes.CDLL(PROBE)


def _run(fn, arg):
    out = (ctypes.c_uint32 * 2)()
    assert fn(arg, out) == 0
    return out[0], (out[1] >> 16, out[1] & 0xFFFF)


def test_division_all_pairs(probe):
    bad, first = _run(probe.probe_div_small, 32768)
    assert bad == 0, f"{bad} quotients differ; first (b, a) = {first}"


def test_magnitude_rounding_all_pairs(probe):
    bad, first = _run(probe.probe_mag_body, 32768)
    assert bad == 0, f"{bad} magnitudes differ; first (gx, gy) = {first}"


@pytest.mark.parametrize("gradient", [3])
def test_harris_norm_division_all_floats(probe, gradient):
    """the Harris score divides by the fixed norm (255 * 2^(g-1) * 3)^4; the reciprocal-and-FMA
    replacement equals IEEE division for every finite float dividend"""
    f = np.float32(255.0) * np.float32(1 << (gradient - 1)) * np.float32(3.0)
    norm = ((f * f) * f) * f
    out = (ctypes.c_uint32 * 2)()
    assert probe.probe_div_const(ctypes.c_float(float(norm)), out) == 0
    assert out[0] == 0, f"{out[0]} quotients differ; first dividend bits {out[1]:#010x}"


def test_dpp_wave_shift_left_lane_map(probe):
    out = (ctypes.c_int32 * 64)()
    assert probe.probe_dpp_wave_shl1(out) == 0
    assert list(out)[:63] == [1001 + i for i in range(63)]


def test_canny_l2_pack_all_integer_floats(probe):
    """the Canny L2 magnitude's square-root-and-round rule (correctly rounded sqrt, nearest-even
    conversion, 16-bit saturation) against the SSE rule restated in double, for every integer-valued
    float below 2^32"""
    out = (ctypes.c_uint32 * 2)()
    assert probe.probe_l2_pack(out) == 0
    assert out[0] == 0, f"{out[0]} values differ; first integer {out[1]}"
