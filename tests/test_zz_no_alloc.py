"""The allocation-free forward 2-D transform (sw_ptospc_noalloc) against sw_ptospc.

sw_ptospc_noalloc runs the same two passes as sw_ptospc with the buffer between them in the library's
fixed per-device scratch array, on device arrays the caller owns.  Two calls in a row, the second on a
smaller grid (the scratch keeps the first call's contents), must each equal sw_ptospc bit for bit; a
grid beyond the scratch and host arrays are refused with SW_ERR_ARG.  The device arrays are allocated
by the test through the HIP runtime that libsw_hip.so is linked against."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
H2D, D2H = 1, 2  # hipMemcpyKind


@pytest.fixture(scope="module")
def hip(sw):
    h = C.CDLL("libamdhip64.so.7")  # already loaded as a dependency of libsw_hip.so: the same runtime
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return h


def dev_array(hip, a):
    p = C.c_void_p()
    assert hip.hipMalloc(C.byref(p), C.c_size_t(a.nbytes)) == 0
    assert hip.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), H2D) == 0
    return p


def noalloc(sw, d, pin, pout):
    sw._lib.sw_ptospc_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    return sw._lib.sw_ptospc_noalloc(d._h, pin, pout)


def test_noalloc_matches_ptospc_two_calls(sw, hip):
    rng = np.random.default_rng(42)
    for ng in (512, 96):  # the second call smaller than the first, and on the block-per-rows passes
        d = sw.D2FFT(ng, ng)
        x = rng.standard_normal(ng * ng)
        ref = np.zeros(ng * ng)
        d.ptospc(x.copy(), ref)
        pin = dev_array(hip, x)
        pout = dev_array(hip, np.full(ng * ng, np.nan))
        try:
            assert noalloc(sw, d, pin, pout) == 0, sw._lib.sw_last_error()
            out = np.empty(ng * ng)
            assert hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), pout, C.c_size_t(out.nbytes), D2H) == 0
            back = np.empty(ng * ng)
            assert hip.hipMemcpy(back.ctypes.data_as(C.c_void_p), pin, C.c_size_t(back.nbytes), D2H) == 0
        finally:
            hip.hipFree(pin)
            hip.hipFree(pout)
        print(f"ng={ng}: max |noalloc - ptospc| = {np.abs(out - ref).max():.1e}", flush=True)
        assert np.array_equal(out, ref), ng
        assert np.array_equal(back, x), ng  # the input is left intact


def test_noalloc_refuses_large_grid_and_host_arrays(sw, hip):
    d = sw.D2FFT(1024, 1024)
    x = np.zeros(1024 * 1024)
    pin = dev_array(hip, x)
    try:
        assert noalloc(sw, d, pin, pin) == -1
        assert b"exceeds its scratch" in sw._lib.sw_last_error()
    finally:
        hip.hipFree(pin)
    d = sw.D2FFT(64, 64)
    a, b = np.zeros(64 * 64), np.zeros(64 * 64)
    assert noalloc(sw, d, a.ctypes.data_as(C.c_void_p), b.ctypes.data_as(C.c_void_p)) == -1
    assert b"device arrays only" in sw._lib.sw_last_error()
