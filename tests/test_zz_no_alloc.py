"""c3d_monitor_noalloc: the field summaries of c3d_monitor computed on device
arrays the caller owns, with the library's static device scratch instead of
allocations.  Checked bitwise against c3d_monitor (same kernels, same order)
on two calls in a row, the second with a smaller grid, through the C ABI."""
import ctypes as C

import numpy as np
import pytest

from channel3d_parallel_amd import Channel, fields
from channel3d_parallel_amd import _lib

pytestmark = pytest.mark.gpu


def _hip():
    h = C.CDLL("libamdhip64.so")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


H2D, D2H = 1, 2


def _device_copy(h, a):
    p = C.c_void_p()
    assert h.hipMalloc(C.byref(p), a.nbytes) == 0
    assert h.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, H2D) == 0
    return p


@pytest.mark.parametrize("sizes", [((66, 33, 34), (34, 17, 18))])
def test_monitor_noalloc_matches_monitor(sizes):
    h = _hip()
    L = _lib.load()
    L.c3d_monitor_noalloc.restype = C.c_int
    L.c3d_monitor_noalloc.argtypes = [C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_void_p, C.c_void_p]
    for call, (nx, ny, nz) in enumerate(sizes):
        print(f"\n[case] monitor_noalloc call {call + 1}: {nx}x{ny}x{nz} faces", flush=True)
        U, V, W = fields.white_faces(nx, ny, nz, seed=40 + call)
        U[3, 4, 5] = np.inf  # a non-finite entry in the interior is counted
        with Channel(nx, ny, nz) as ch:
            ch.set_fields(U, V, W)
            ref = ch.monitor()
        dU, dV, dW = (_device_copy(h, np.asfortranarray(a)) for a in (U, V, W))
        dout = C.c_void_p()
        assert h.hipMalloc(C.byref(dout), 7 * 8) == 0
        try:
            rc = L.c3d_monitor_noalloc(nx, ny, nz, dU, dV, dW, dout, None)
            assert rc == 0, rc
            assert h.hipDeviceSynchronize() == 0
            out = np.zeros(7)
            assert h.hipMemcpy(out.ctypes.data_as(C.c_void_p), dout, out.nbytes, D2H) == 0
        finally:
            for p in (dU, dV, dW, dout):
                h.hipFree(p)
        want = np.array([ref[k] for k in ("maxU", "maxV", "maxW", "meanU", "meanV", "meanW",
                                          "nonfinite")])
        print(f"  noalloc {out}\n  monitor {want}")
        assert np.array_equal(out, want)


def test_monitor_noalloc_rejects_bad_sizes():
    L = _lib.load()
    L.c3d_monitor_noalloc.restype = C.c_int
    L.c3d_monitor_noalloc.argtypes = [C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_void_p, C.c_void_p]
    assert L.c3d_monitor_noalloc(2, 17, 18, None, None, None, None, None) == -1
