"""ctypes front-end of the oracle (TEST INFRASTRUCTURE: the checker, never the product).

* ``Oracle``: the C restatement, oracle/liboracle.so (built by ``make -C oracle``).
* ``Reference``: the reference's own CPU engine driven through oracle/_ref/libref_harness.so,
  present only where oracle/_ref was built from the reference sources.

Every function takes and returns packed numpy arrays (row stride = width * bytes per pixel).
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libref_harness.so")


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def build_oracle():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


class Oracle:
    def __init__(self):
        if not os.path.exists(ORACLE_SO):
            build_oracle()
        self.lib = ctypes.CDLL(ORACLE_SO)
        self.lib.oracle_magnitude_px.restype = ctypes.c_int16
        self.lib.oracle_phase_px.restype = ctypes.c_uint8

    def filter3x3(self, which, src):
        h, w = src.shape
        dst = np.zeros((h, w), np.uint8)
        self.lib.oracle_filter3x3(which, w, h, _p(np.ascontiguousarray(src)), _p(dst))
        return dst

    def sobel(self, src):
        h, w = src.shape
        gx = np.zeros((h, w), np.int16)
        gy = np.zeros((h, w), np.int16)
        self.lib.oracle_sobel_node(w, h, _p(np.ascontiguousarray(src)), _p(gx), _p(gy))
        return gx, gy

    def magnitude(self, gx, gy):
        h, w = gx.shape
        out = np.zeros((h, w), np.int16)
        self.lib.oracle_magnitude(w, h, _p(out), 2 * w, _p(np.ascontiguousarray(gx)), 2 * w,
                                  _p(np.ascontiguousarray(gy)), 2 * w)
        return out

    def phase(self, gx, gy):
        h, w = gx.shape
        out = np.zeros((h, w), np.uint8)
        self.lib.oracle_phase(w, h, _p(out), w, _p(np.ascontiguousarray(gx)), 2 * w,
                              _p(np.ascontiguousarray(gy)), 2 * w)
        return out

    def chain1(self, src):
        h, w = src.shape
        gx = np.zeros((h, w), np.int16); gy = np.zeros((h, w), np.int16)
        mag = np.zeros((h, w), np.int16); ph = np.zeros((h, w), np.uint8)
        self.lib.oracle_chain1(w, h, _p(np.ascontiguousarray(src)), _p(gx), _p(gy), _p(mag), _p(ph))
        return gx, gy, mag, ph


class Reference:
    """The reference CPU engine; raises FileNotFoundError when oracle/_ref was not built."""

    def __init__(self):
        if not os.path.exists(REF_SO):
            raise FileNotFoundError(REF_SO)
        self.lib = ctypes.CDLL(REF_SO)

    def _ok(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"reference {what} failed: {rc}")

    def filter3x3(self, which, src):
        h, w = src.shape
        dst = np.zeros((h, w), np.uint8)
        self._ok(self.lib.ref_filter3x3(which, w, h, _p(np.ascontiguousarray(src)), _p(dst)), "filter3x3")
        return dst

    def sobel(self, src):
        h, w = src.shape
        gx = np.zeros((h, w), np.int16)
        gy = np.zeros((h, w), np.int16)
        self._ok(self.lib.ref_sobel3x3(w, h, _p(np.ascontiguousarray(src)), _p(gx), _p(gy)), "sobel")
        return gx, gy

    def magnitude_phase(self, gx, gy):
        h, w = gx.shape
        mag = np.zeros((h, w), np.int16)
        ph = np.zeros((h, w), np.uint8)
        self._ok(self.lib.ref_magnitude_phase(w, h, _p(np.ascontiguousarray(gx)), _p(np.ascontiguousarray(gy)),
                                              _p(mag), _p(ph)), "magnitude_phase")
        return mag, ph

    def chain1(self, src, time_iters=0):
        h, w = src.shape
        gx = np.zeros((h, w), np.int16); gy = np.zeros((h, w), np.int16)
        mag = np.zeros((h, w), np.int16); ph = np.zeros((h, w), np.uint8)
        sec = ctypes.c_double(0.0)
        self._ok(self.lib.ref_chain1(w, h, _p(np.ascontiguousarray(src)), _p(gx), _p(gy), _p(mag), _p(ph),
                                     time_iters, ctypes.byref(sec)), "chain1")
        if time_iters:
            return (gx, gy, mag, ph), sec.value
        return gx, gy, mag, ph
