"""CPU-side checks of the C-ABI library: it loads, exports every function
include/sipg_hip.h declares, and its host-only logic (forcing draw,
argument validation that needs no device) behaves.  No device compute."""
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "sipg_hip.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(sipg_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_the_core_api():
    names = declared_functions()
    for must in ("sipg_op_create", "sipg_apply", "sipg_cg", "sipg_dot", "sipg_axpy",
                 "sipg_xpby", "sipg_alpha_update", "sipg_op_destroy", "sipg_strerror"):
        assert must in names


def test_library_exports_every_declared_symbol(sipg):
    L = sipg.lib()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_strerror_and_max_degree(sipg):
    assert sipg.strerror(0) == "ok"
    assert "overlap" in sipg.strerror(2)
    assert sipg.max_degree() >= 15


@pytest.mark.parametrize("seed,kmin,kmax,mmax", [(0, 4, 4, 4), (1, 1, 4, 4), (7, 1, 8, 12), (2**40 + 3, 1, 4, 4)])
def test_forcing_draw_matches_mt19937_64(sipg, seed, kmin, kmax, mmax):
    from mt19937_64 import forcing_draw
    a, m, n = sipg.forcing_draw(seed, kmin, kmax, mmax)
    ra, rm, rn = forcing_draw(seed, kmin, kmax, mmax)
    assert len(a) == len(ra)
    assert np.array_equal(a, np.array(ra)) and list(m) == rm and list(n) == rn
    assert np.all(np.abs(a) <= 1) and np.all((m >= 1) & (m <= mmax))


def test_op_create_rejects_bad_arguments(sipg):
    import ctypes
    h = ctypes.c_void_p()
    L = sipg.lib()
    assert L.sipg_op_create(ctypes.byref(h), 0, 4, 3, 1.0, 0, None) == sipg.ERR_INVALID
    assert L.sipg_op_create(ctypes.byref(h), 4, 4, 0, 1.0, 0, None) == sipg.ERR_INVALID
    assert L.sipg_op_create(ctypes.byref(h), 4, 4, 3, 1.0, 7, None) == sipg.ERR_INVALID
    assert L.sipg_op_create(ctypes.byref(h), 4, 4, 3, float("nan"), 0, None) == sipg.ERR_INVALID
    assert L.sipg_op_create(ctypes.byref(h), 4, 4, sipg.max_degree() + 1, 1.0, 0, None) == sipg.ERR_UNSUPPORTED
