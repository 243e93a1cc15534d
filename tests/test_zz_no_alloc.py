"""fw_cc_pair_noalloc against fw_cc_pair through the C ABI: the same images give the same bits
(derivative fields and values), on two calls in a row, the second on a smaller grid (so the static
scratch keeps the first call's partials when the second starts). The caller's buffers come from
hipMalloc of the HIP runtime the library itself loaded; the entry point allocates nothing."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from fastwarp_amd import backend as fb
from tests import cases
from tests.test_gpu_cc_precision import level_images

pytestmark = pytest.mark.gpu


def hip():
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


H2D, D2H = 1, 2


class DeviceArray:
    def __init__(self, h, nbytes):
        self.h = h
        self.p = C.c_void_p()
        assert h.hipMalloc(C.byref(self.p), nbytes) == 0
        self.nbytes = nbytes

    def put(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        assert a.nbytes == self.nbytes
        assert self.h.hipMemcpy(self.p, a.ctypes.data_as(C.c_void_p), a.nbytes, H2D) == 0

    def get(self, dtype, count) -> np.ndarray:
        out = np.empty(count, dtype=dtype)
        assert out.nbytes == self.nbytes
        assert self.h.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        if self.p:
            self.h.hipFree(self.p)
            self.p = C.c_void_p()


def noalloc_pair(h, g, f, m, radius):
    n = g.n
    bufs = [DeviceArray(h, 4 * n), DeviceArray(h, 4 * n), DeviceArray(h, 12 * n), DeviceArray(h, 12 * n), DeviceArray(h, 16)]
    try:
        bufs[0].put(f.astype(np.float32))
        bufs[1].put(m.astype(np.float32))
        st = fb.lib().fw_cc_pair_noalloc(C.byref(g), bufs[0].p, bufs[1].p, radius, bufs[2].p, bufs[3].p, bufs[4].p, None)
        assert st == fb.FW_OK, fb.lib().fw_last_error(None)
        assert h.hipDeviceSynchronize() == 0
        d0 = bufs[2].get(np.float32, 3 * n).reshape(3, n).T.astype(np.float64)
        d1 = bufs[3].get(np.float32, 3 * n).reshape(3, n).T.astype(np.float64)
        values = bufs[4].get(np.float64, 2)
        return values, d0, d1
    finally:
        for b in bufs:
            b.free()


def test_noalloc_cc_pair_matches_fw_cc_pair_on_two_calls(backend):
    h = hip()
    cases_ = [level_images((182, 218, 182), 1.0, 2, 1, 0.0), level_images((96, 112, 96), 2.0, 1, 2, 1.0)]
    for g, f, m in cases_:
        values, d0, d1 = noalloc_pair(h, g, f, m, 4)
        (v0, e0), (v1, e1) = backend.cc_value_derivative_pair(backend.upload_image(g, f), backend.upload_image(g, m), 4)
        assert values[0] == v0 and values[1] == v1, (values, v0, v1)
        assert np.array_equal(d0, backend.download_field(e0))
        assert np.array_equal(d1, backend.download_field(e1))


def test_noalloc_cc_pair_rejects_what_its_scratch_cannot_hold(backend):
    h = hip()
    tiny = DeviceArray(h, 64)
    try:
        big = cases.grid((1024, 1024, 512))
        st = fb.lib().fw_cc_pair_noalloc(C.byref(big), tiny.p, tiny.p, 4, tiny.p, tiny.p, tiny.p, None)
        assert st == fb.FW_ERR_INVALID and b"static scratch" in fb.lib().fw_last_error(None)
        small = cases.grid((8, 8, 8))
        st = fb.lib().fw_cc_pair_noalloc(C.byref(small), tiny.p, tiny.p, 9, tiny.p, tiny.p, tiny.p, None)
        assert st == fb.FW_ERR_INVALID and b"outside 1..=8" in fb.lib().fw_last_error(None)
    finally:
        tiny.free()
