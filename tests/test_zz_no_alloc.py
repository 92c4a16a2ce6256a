"""cx_hilbert_order_noalloc against cx_hilbert_order through the C-ABI: the
allocation-free entry point (work buffers in one fixed device array per GPU)
must give the same permutation bit for bit, on two calls in a row, the second
with a smaller input, and reject inputs beyond its fixed maximum."""

import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAX_POINTS = 1 << 18
H2D, D2H = 1, 2


def _hip_runtime():
    # the HIP runtime the library itself links (the test allocates its own
    # device buffers; only the entry point under test must not)
    rt = C.CDLL("/opt/rocm/lib/libamdhip64.so.7")
    rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    rt.hipFree.argtypes = [C.c_void_p]
    rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    rt.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    return rt


def _dev(rt, nbytes):
    p = C.c_void_p()
    assert rt.hipMalloc(C.byref(p), nbytes) == 0
    return p


def _device_order(pts):
    from covertreex_amd import _ffi

    rt = _hip_runtime()
    x = np.ascontiguousarray(pts, np.float64)
    n, d = x.shape
    dx, do = _dev(rt, x.nbytes), _dev(rt, 8 * n)
    try:
        assert rt.hipMemcpy(dx, x.ctypes.data, x.nbytes, H2D) == 0
        assert rt.hipMemset(do, 0xff, 8 * n) == 0
        assert rt.hipDeviceSynchronize() == 0
        st = _ffi.load().cx_hilbert_order_noalloc(dx, n, d, do, None)
        assert st == _ffi.CX_OK, _ffi.load().cx_last_error()
        assert rt.hipDeviceSynchronize() == 0
        out = np.empty(n, np.int64)
        assert rt.hipMemcpy(out.ctypes.data, do, 8 * n, D2H) == 0
        return out
    finally:
        rt.hipFree(dx)
        rt.hipFree(do)


def test_noalloc_hilbert_matches_allocating_entry_point(hip):
    from covertreex_amd.order import hilbert_permutation

    rng = np.random.default_rng(12)
    big = rng.standard_normal((200_000, 3))
    big[1000:2000] = big[:1000]                  # duplicate rows: ties keep index order
    small = rng.uniform(-1.0, 1.0, (5_000, 8))
    for pts in (big, small):                     # the second call reuses the scratch array
        got = _device_order(pts)
        want = hilbert_permutation(pts)
        assert np.array_equal(got, want)


def test_noalloc_hilbert_rejects_beyond_its_maximum(hip):
    # refused on the host before any device work: the pointers are never read
    from covertreex_amd import _ffi

    rt = _hip_runtime()
    dx, do = _dev(rt, 16), _dev(rt, 16)
    try:
        st = _ffi.load().cx_hilbert_order_noalloc(dx, MAX_POINTS + 1, 2, do, None)
        assert st == _ffi.CX_ERR_INVALID
        assert b"262144" in _ffi.load().cx_last_error()
    finally:
        rt.hipFree(dx)
        rt.hipFree(do)
