"""sipg_vec_dot_noalloc (scratch in a fixed device array) against
sipg_vec_dot (scratch allocated per call), through the C-ABI: the same
kernels in the same order, so the results agree bit for bit.  Two calls in
a row per dtype, the second on a smaller input, so the second reuses the
scratch the first left behind."""
import ctypes

import numpy as np
import pytest


def _fn(lib, name):
    f = getattr(lib, name)
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_int, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
                  ctypes.POINTER(ctypes.c_double), ctypes.c_void_p]
    return f


@pytest.mark.gpu
def test_vec_dot_noalloc_matches_vec_dot(sipg):
    lib = sipg.lib()
    sipg.check(lib.sipg_set_device(0), "set_device")
    ref, noalloc = _fn(lib, "sipg_vec_dot"), _fn(lib, "sipg_vec_dot_noalloc")
    rng = np.random.default_rng(7)
    for dtype, np_t in ((sipg.F64, np.float64), (sipg.F32, np.float32)):
        # 2.5M values need the full 1024 block sums; 4097 need only a few,
        # so the second call leaves stale sums of the first in the scratch
        for n in (2_500_003, 4097, 0):
            a = rng.uniform(-1, 1, max(n, 1)).astype(np_t)
            b = rng.uniform(-1, 1, max(n, 1)).astype(np_t)
            da, db = sipg.DeviceBuffer(a.nbytes), sipg.DeviceBuffer(b.nbytes)
            da.upload(a)
            db.upload(b)
            r1, r2 = ctypes.c_double(np.nan), ctypes.c_double(np.nan)
            sipg.check(ref(dtype, n, da.ptr, db.ptr, ctypes.byref(r1), None), "vec_dot")
            sipg.check(noalloc(dtype, n, da.ptr, db.ptr, ctypes.byref(r2), None), "vec_dot_noalloc")
            assert r2.value == r1.value, (dtype, n, r1.value, r2.value)
            exact = float(np.dot(a[:n].astype(np.float64), b[:n].astype(np.float64)))
            assert abs(r2.value - exact) <= 1e-12 * max(1.0, float(np.abs(a[:n]).sum())), (dtype, n)


@pytest.mark.gpu
def test_vec_dot_noalloc_rejects_bad_arguments(sipg):
    lib = sipg.lib()
    noalloc = _fn(lib, "sipg_vec_dot_noalloc")
    r = ctypes.c_double(0.0)
    buf = sipg.DeviceBuffer(64)
    assert noalloc(sipg.F64, -1, buf.ptr, buf.ptr, ctypes.byref(r), None) == sipg.ERR_INVALID
    assert noalloc(7, 4, buf.ptr, buf.ptr, ctypes.byref(r), None) == sipg.ERR_INVALID
    assert noalloc(sipg.F64, 4, None, buf.ptr, ctypes.byref(r), None) == sipg.ERR_INVALID
