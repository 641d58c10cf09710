"""The C ABI library loads and exports every entry point include/*.h declares (CPU: no compute)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mvx_\w+)\s*\(", text)))


def test_kernel_library_exports_every_declared_symbol():
    import mivisionx_amd
    lib = mivisionx_amd.kernels()
    names = _declared("mvx_kernels.h")
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    lib.mvx_version.restype = ctypes.c_char_p
    assert b"gfx950" in lib.mvx_version()


def test_workspace_size_queries_are_host_only():
    import mivisionx_amd
    lib = mivisionx_amd.kernels()
    lib.mvx_NonMaxSupp_XY_ANY_3x3_workspace_size.restype = ctypes.c_size_t
    lib.mvx_HarrisMergeSortAndPick_XY_XYS_workspace_size.restype = ctypes.c_size_t
    assert lib.mvx_NonMaxSupp_XY_ANY_3x3_workspace_size(3840, 2160) >= 4 * 15 * 2158
    assert lib.mvx_HarrisMergeSortAndPick_XY_XYS_workspace_size(1000) >= 8 * 1024 + 8 * 2048


@pytest.mark.parametrize("bad", ["stride", "align"])
def test_parameter_validation_needs_no_gpu(bad):
    """invalid strides are rejected on the host before any launch"""
    import mivisionx_amd
    lib = mivisionx_amd.kernels()
    p = ctypes.c_void_p(16 * 1024)      # never dereferenced: validation fails first
    q = ctypes.c_void_p(16 * 1024 + 3)
    if bad == "stride":
        rc = lib.mvx_Gaussian_U8_U8_3x3(None, 64, 8, p, 40, p, 64)
    else:
        rc = lib.mvx_Gaussian_U8_U8_3x3(None, 64, 8, q, 64, p, 64)
    assert rc == -10


@pytest.mark.parametrize("name", ["mvx_Remap_U8_U8_Bilinear", "mvx_Remap_U8_U8_Nearest", "mvx_Remap_U8_U8_Nearest_Constant"])
def test_remap_source_buffer_size_is_checked(name):
    """the reference's srcImageBufferSize argument: a buffer smaller than the source plane is rejected on
    the host before any launch"""
    import mivisionx_amd
    fn = getattr(mivisionx_amd.kernels(), name)
    p = ctypes.c_void_p(16 * 1024)      # never dereferenced: validation fails first
    args = [None, 64, 8, p, 64, 64, 8, p, 64, ctypes.c_uint32(64 * 7 + 63), p, 256]
    if name.endswith("Constant"):
        args.append(ctypes.c_uint8(0))
    assert fn(*args) == -10
