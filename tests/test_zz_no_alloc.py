"""The allocation-free entry point pp_ten_masked_logdot_noalloc (device
buffers in and out, the per-site util vector in the library's static device
array) against pp_ten_masked_logdot, the twin of ten_masked_logdot_c
(linear_c.c:465-539) it is a variant of: bit-equal results on two calls in a
row, the second shorter, and a clear error beyond the stated maximum.

The test's own device buffers come from the HIP runtime the library is linked
against (a second runtime in the process, such as torch's, would not see the
device once the library has initialised it)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAX_SITES = 1 << 18
H2D, D2H = 1, 2


class _Dev:
    """A device buffer from the library's HIP runtime."""

    def __init__(self, hip, nbytes):
        self.hip = hip
        self.p = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), C.c_size_t(max(nbytes, 8))) == 0

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert self.hip.hipMemcpy(self.p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), H2D) == 0

    def get(self, a):
        assert self.hip.hipMemcpy(a.ctypes.data_as(C.c_void_p), self.p, C.c_size_t(a.nbytes), D2H) == 0

    def free(self):
        self.hip.hipFree(self.p)


def test_noalloc_masked_logdot_equals_the_twin():
    from pplacer_amd import _lib, linear
    print("\n[gpu case] noalloc masked logdot", flush=True)
    L = _lib.lib()
    hip = C.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    rng = np.random.default_rng(5)
    for R, N, S in ((4, 3000, 4), (4, 700, 20)):  # the second call shorter
        x, y = rng.random((R, N, S)), rng.random((R, N, S))
        mask = (rng.random(N) < 0.7).astype(np.uint16)
        want = linear.ten_masked_logdot(x, y, mask)
        bufs = [_Dev(hip, a.nbytes) for a in (x, y, mask)]
        for b, a in zip(bufs, (x, y, mask)):
            b.put(a)
        dout = _Dev(hip, 8)
        dout.put(np.array([np.nan]))
        rc = L.pp_ten_masked_logdot_noalloc(bufs[0].p, bufs[1].p, bufs[2].p, R, N, S, dout.p, None)
        assert rc == 0, L.pp_last_error()
        assert hip.hipDeviceSynchronize() == 0
        got = np.empty(1)
        dout.get(got)
        assert got[0] == want, (R, N, S, got[0], want)
        for b in bufs + [dout]:
            b.free()
    # beyond the stated maximum: an error, nothing launched
    rc = L.pp_ten_masked_logdot_noalloc(None, None, None, 1, 10, 4, None, None)
    assert rc == 1  # null buffers
    d = _Dev(hip, 64)
    rc = L.pp_ten_masked_logdot_noalloc(d.p, d.p, d.p, 1, MAX_SITES + 1, 4, d.p, None)
    assert rc == 1
    assert b"exceeds" in L.pp_last_error()
    d.free()
