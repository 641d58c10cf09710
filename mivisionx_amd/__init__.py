"""mivisionx_amd -- MI355X-native OpenVX vision-node hot path.

The product is two native libraries built in-tree:

* ``libmvx_kernels.so``: hand-written gfx950 HIP kernels behind the C ABI in
  ``include/mvx_kernels.h`` (one entry point per reference ``HipExec_*`` launcher);
* ``libmvx_openvx.so``: the C++ graph engine exposing the ``vx*``/``vxu*`` API.

This Python package only loads them (ctypes) and moves data for tests and benchmarks; PyTorch
provides device memory and streams.  Nothing here computes a node on the CPU: when a library is
missing the loaders raise.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
KERNELS_SO = os.path.join(_HERE, "libmvx_kernels.so")

_kernels = None


class NativeLibraryMissing(RuntimeError):
    pass


def kernels():
    """The loaded libmvx_kernels.so (raises NativeLibraryMissing when it is not built)."""
    global _kernels
    if _kernels is None:
        if not os.path.exists(KERNELS_SO):
            raise NativeLibraryMissing(
                f"{KERNELS_SO} is not built; run `python -c 'import __graft_entry__ as g; g.build()'`")
        _kernels = ctypes.CDLL(KERNELS_SO)
        _kernels.mvx_version.restype = ctypes.c_char_p
    return _kernels
