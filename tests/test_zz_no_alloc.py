"""dcrf_less_confident_pixels_noalloc (no device allocation, fixed scratch)
through the C ABI against dcrf_less_confident_pixels on the same device
buffers: two calls in a row, the second with a smaller field, give the same
indices and counts; an empty field, the size limit and host pointers are
handled without a launch."""

import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

L = 21


def _field(N, seed):
    rng = np.random.default_rng(seed)
    Q = rng.random((N, L)).astype(np.float32)
    Q /= Q.sum(1, keepdims=True)
    conf = rng.random(N) < 0.6  # most pixels confident
    Q[conf, rng.integers(0, L, conf.sum())] = 0.995
    return Q


def _both(lib, D, Q, tol):
    N = Q.shape[0]
    q = D.DeviceArray.from_numpy(Q)
    i1 = D.DeviceArray((max(N, 1),), np.int32)
    i2 = D.DeviceArray((max(N, 1),), np.int32)
    n1, n2 = C.c_int(-1), C.c_int(-1)
    assert lib.dcrf_less_confident_pixels(q.ptr, N, L, C.c_float(tol), i1.ptr, C.byref(n1), 1) == 0
    rc = lib.dcrf_less_confident_pixels_noalloc(q.ptr, N, L, C.c_float(tol), i2.ptr, C.byref(n2))
    assert rc == 0, lib.dcrf_last_error()
    return i1.numpy()[: n1.value], n1.value, i2.numpy()[: n2.value], n2.value


def test_noalloc_matches_allocating_entry_point():
    import densecrf_ho_amd as D
    from densecrf_ho_amd import _lib
    lib = _lib.load()
    # 1920x1080 then a smaller, ragged field (not a multiple of the tile)
    for N, seed, tol in ((1920 * 1080, 1, 0.99), (123457, 2, 0.99)):
        print("less_confident_pixels_noalloc N", N, flush=True)
        Q = _field(N, seed)
        a, na, b, nb = _both(lib, D, Q, tol)
        want = np.nonzero(Q.max(1) <= np.float32(tol))[0]
        assert na == nb == len(want) > 0
        np.testing.assert_array_equal(a, want)
        np.testing.assert_array_equal(b, want)


def test_noalloc_edge_cases():
    import densecrf_ho_amd as D
    from densecrf_ho_amd import _lib
    lib = _lib.load()
    n = C.c_int(-1)
    # empty field: count 0, nothing launched
    assert lib.dcrf_less_confident_pixels_noalloc(None, 0, L, C.c_float(0.9), None, C.byref(n)) == 0
    assert n.value == 0
    assert lib.dcrf_less_confident_pixels(None, 0, L, C.c_float(0.9), None, C.byref(n), 1) == 0 and n.value == 0
    q = D.DeviceArray((64, L))
    idx = D.DeviceArray((64,), np.int32)
    # beyond the fixed scratch: refused before any launch
    rc = lib.dcrf_less_confident_pixels_noalloc(q.ptr, 16777217, L, C.c_float(0.9), idx.ptr, C.byref(n))
    assert rc == 1 and b"16,777,216" in lib.dcrf_last_error()
    # host memory is refused, not read
    host = np.zeros((64, L), np.float32)
    rc = lib.dcrf_less_confident_pixels_noalloc(host.ctypes.data_as(C.c_void_p), 64, L, C.c_float(0.9), idx.ptr,
                                                C.byref(n))
    assert rc == 1 and b"device pointer" in lib.dcrf_last_error()
    # a single pixel, confident and not
    Q = np.full((1, L), 1.0 / L, np.float32)
    a, na, b, nb = _both(lib, D, Q, 0.5)
    assert na == nb == 1 and a[0] == b[0] == 0
    Q[0, 3] = 0.99
    a, na, b, nb = _both(lib, D, Q, 0.5)
    assert na == nb == 0
