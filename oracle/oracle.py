"""ctypes front end of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: the checker for the HIP product path and the CPU
baseline timed beside it.  Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg may import this module.

Every function forwards to the C restatement in sipg_oracle.c, which cites
the reference file:line it follows.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libref_basis.so")

_lib = None
_ref = None

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int)


class Problem(ctypes.Structure):
    """orc_problem: kind 1..5 = reference solutions A..E, 0 = Fourier modes."""
    _fields_ = [("kind", ctypes.c_int), ("nmodes", ctypes.c_int),
                ("a", _dp), ("m", _ip), ("n", _ip)]


def build():
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        L.orc_legendre.argtypes = [ctypes.c_int, ctypes.c_double, _dp, _dp]
        L.orc_gll.argtypes = [ctypes.c_int, _dp, _dp]
        L.orc_dphi.argtypes = [ctypes.c_int, _dp, _dp]
        L.orc_k1.argtypes = [ctypes.c_int, _dp, _dp, _dp]
        L.orc_apply.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp, _dp]
        L.orc_apply_f32.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float, _fp, _fp]
        L.orc_rhs.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.POINTER(Problem), _dp]
        L.orc_cg.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp, _dp,
                             ctypes.c_double, ctypes.c_int, ctypes.c_int, _dp]
        L.orc_cg.restype = ctypes.c_int
        L.orc_err_norms.argtypes = [ctypes.c_int, ctypes.c_int, _dp, ctypes.POINTER(Problem), _dp]
        L.orc_solve.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.POINTER(Problem),
                                ctypes.c_double, ctypes.c_int, _dp, _dp]
        L.orc_solve.restype = ctypes.c_int
        L.orc_mesh_coords.argtypes = [ctypes.c_int, _dp, _dp]
        L.orc_print_result.argtypes = [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_char_p]
        L.orc_print_result.restype = ctypes.c_int
        for fn in ("orc_u_ex", "orc_f", "orc_dx_u_ex", "orc_dy_u_ex"):
            getattr(L, fn).argtypes = [ctypes.POINTER(Problem), ctypes.c_double, ctypes.c_double]
            getattr(L, fn).restype = ctypes.c_double
        L.orc_num_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def legendre(n, z):
    P, dP = ctypes.c_double(), ctypes.c_double()
    lib().orc_legendre(n, z, ctypes.byref(P), ctypes.byref(dP))
    return P.value, dP.value


def gll(p):
    z = np.zeros(p + 1)
    w = np.zeros(p + 1)
    rc = lib().orc_gll(p, _d(z), _d(w))
    assert rc == 0
    return z, w


def dphi(p):
    z, _ = gll(p)
    D = np.zeros((p + 1) * (p + 1))
    lib().orc_dphi(p, _d(z), _d(D))
    return D.reshape(p + 1, p + 1)   # D[mode, node]


def k1(p):
    z, w = gll(p)
    D = dphi(p).ravel().copy()
    K = np.zeros((p + 1) * (p + 1))
    lib().orc_k1(p, _d(w), _d(D), _d(K))
    return K.reshape(p + 1, p + 1)


def apply(dimx, dimy, p, pen, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty_like(x)
    lib().orc_apply(dimx, dimy, p, pen, _d(x), _d(y))
    return y


def apply_f32(dimx, dimy, p, pen, x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.empty_like(x)
    lib().orc_apply_f32(dimx, dimy, p, pen, x.ctypes.data_as(_fp), y.ctypes.data_as(_fp))
    return y


class _Prob:
    """Keeps the ctypes Problem and its arrays alive together."""

    def __init__(self, kind, modes=None):
        self.a = np.zeros(1)
        self.m = np.zeros(1, dtype=np.int32)
        self.n = np.zeros(1, dtype=np.int32)
        nm = 0
        if modes is not None:
            a, m, n = modes
            self.a = np.ascontiguousarray(a, dtype=np.float64)
            self.m = np.ascontiguousarray(m, dtype=np.int32)
            self.n = np.ascontiguousarray(n, dtype=np.int32)
            nm = len(self.a)
        self.s = Problem(kind, nm, _d(self.a), self.m.ctypes.data_as(_ip), self.n.ctypes.data_as(_ip))

    def ref(self):
        return ctypes.byref(self.s)


def problem(kind, modes=None):
    return _Prob(kind, modes)


def rhs(dim, p, pen, prob):
    b = np.zeros(dim * dim * (p + 1) ** 2)
    lib().orc_rhs(dim, p, pen, prob.ref(), _d(b))
    return b


def cg(dimx, dimy, p, pen, b, x0=None, tol=1e-10, tol_mode=1, maxit=0):
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros_like(b) if x0 is None else np.array(x0, dtype=np.float64)
    rr = ctypes.c_double()
    it = lib().orc_cg(dimx, dimy, p, pen, _d(b), _d(x), tol, tol_mode, maxit, ctypes.byref(rr))
    return x, it, rr.value


def err_norms(dim, p, u, prob):
    u = np.ascontiguousarray(u, dtype=np.float64)
    out = np.zeros(3)
    lib().orc_err_norms(dim, p, _d(u), prob.ref(), _d(out))
    return {"max": out[0], "L2": out[1], "H1": out[2]}


def print_result(dim, p, u, path):
    """print_result's text for u (reference order) written to path."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    if lib().orc_print_result(dim, p, _d(u), os.fsencode(path)) != 0:
        raise OSError(f"cannot write {path}")


def solve(dim, p, pen, prob, tol=1e-14, tol_mode=0):
    u = np.zeros(dim * dim * (p + 1) ** 2)
    out = np.zeros(3)
    it = lib().orc_solve(dim, p, pen, prob.ref(), tol, tol_mode, _d(u), _d(out))
    return u, it, {"max": out[0], "L2": out[1], "H1": out[2]}


def mesh_coords(dim):
    xe = np.zeros(dim * dim)
    ye = np.zeros(dim * dim)
    lib().orc_mesh_coords(dim, _d(xe), _d(ye))
    return xe, ye


def num_threads():
    return lib().orc_num_threads()


def ref_basis_available():
    return os.path.exists(REF_LIB)


def ref_basis(p):
    """The reference's own host numerics (LGL_quadrature_table, sem_function,
    laplace_1d), compiled from its sources into oracle/_ref by `make ref`."""
    global _ref
    if _ref is None:
        _ref = ctypes.CDLL(REF_LIB)
        _ref.ref_basis.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp]
        _ref.ref_basis.restype = ctypes.c_int
    z = np.zeros(p + 1)
    w = np.zeros(p + 1)
    D = np.zeros((p + 1) ** 2)
    K = np.zeros((p + 1) ** 2)
    rc = _ref.ref_basis(p, _d(z), _d(w), _d(D), _d(K))
    assert rc == 0
    return z, w, D.reshape(p + 1, p + 1), K.reshape(p + 1, p + 1)
