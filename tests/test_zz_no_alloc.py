"""mcs_findpeaks_noalloc: the peak finder of mcs_findpeaks with its workspace in the library's
static device scratch (no allocation).  Its output equals mcs_findpeaks' on the same field, on two
calls in a row (the second on a smaller grid), and sizes beyond the scratch are refused before any
launch."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _bind(mcs):
    f = mcs.lib().mcs_findpeaks_noalloc
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p,
                  ctypes.c_int64, ctypes.POINTER(ctypes.c_int64), ctypes.c_int]
    return f


def _noalloc(mcs, field, ths, cap, with_values=True):
    f = _bind(mcs)
    field = np.ascontiguousarray(field)
    d = mcs.to_device(field)
    di = mcs.DeviceBuffer(max(cap, 1) * 8)
    dv = mcs.DeviceBuffer(max(cap, 1) * 8) if with_values else None
    cnt = ctypes.c_int64(-1)
    ext = (ctypes.c_int * 3)(*([field.shape[::-1][i] if i < field.ndim else 1 for i in range(3)]))
    rc = f(d.ptr, ext, ths, di.ptr, dv.ptr if dv else None, cap, ctypes.byref(cnt), 32 if field.dtype == np.float32 else 64)
    n = min(cnt.value, cap)
    idx = di.download((max(n, 0),), np.int64) if n > 0 else np.empty(0, np.int64)
    vals = dv.download((max(n, 0),), np.float64) if (dv and n > 0) else np.empty(0)
    return rc, cnt.value, idx, vals


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_noalloc_matches_findpeaks_on_two_calls(mcs, dt):
    a = np.random.default_rng(11).standard_normal((64, 96, 130)).astype(dt)
    b = np.random.default_rng(12).standard_normal((40, 72)).astype(dt)  # smaller, 2D
    for field, ths in ((a, 0.2), (b, -10.0)):
        want_idx, want_vals = mcs.findpeaks(field, ths, with_values=True)
        rc, n, idx, vals = _noalloc(mcs, field, ths, field.size // 4 + 8)
        assert rc == mcs.OK and n == want_idx.size > 0
        assert np.array_equal(idx, want_idx) and np.array_equal(vals, want_vals)


def test_noalloc_capacity_and_limits(mcs):
    a = np.random.default_rng(3).standard_normal((32, 32, 64)).astype(np.float32)
    want = mcs.findpeaks(a, -10.0)
    rc, n, idx, _ = _noalloc(mcs, a, -10.0, 5, with_values=False)
    assert rc == mcs.ERR_CAPACITY and n == want.size and np.array_equal(idx, want[:5])
    # a grid whose peak mask exceeds the scratch is refused before anything is launched
    f = _bind(mcs)
    d = mcs.to_device(a)
    di = mcs.DeviceBuffer(64)
    cnt = ctypes.c_int64(-1)
    rc = f(d.ptr, (ctypes.c_int * 3)(2048, 2048, 512), 0.0, di.ptr, None, 8, ctypes.byref(cnt), 32)
    assert rc == mcs.ERR_UNSUPPORTED and b"scratch" in mcs.lib().mcs_last_error()
