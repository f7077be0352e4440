"""Python front end of libsipg_hip.so (ctypes over the C-ABI in
include/sipg_hip.h).

This module holds no numerics: every operation is a call into the HIP
library.  If the library is not built, loading fails loudly
(`SipgLibraryMissing`); there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "lib", "libsipg_hip.so")

F64, F32 = 0, 1
TOL_ABS, TOL_REL = 0, 1
METHOD_CG, METHOD_PCG = 0, 1
KERNEL_AUTO, KERNEL_VALU, KERNEL_MFMA, KERNEL_WAVE = 0, 1, 2, 3

OK = 0
ERR_INVALID = 1
ERR_ALIAS = 2
ERR_UNSUPPORTED = 3
ERR_HIP = 4
ERR_NOMEM = 5
NOT_CONVERGED = 6
ERR_NONFINITE = 7
ERR_BREAKDOWN = 8
ERR_BUDGET = 9
ERR_CANCELLED = 10
ERR_NOT_FOUND = 11


class SipgLibraryMissing(RuntimeError):
    pass


class SipgError(RuntimeError):
    def __init__(self, status, what=""):
        self.status = status
        super().__init__(f"{what}: {strerror(status)} (status {status})")


class CgOpts(ctypes.Structure):
    _fields_ = [("tol", ctypes.c_double), ("tol_mode", ctypes.c_int), ("maxit", ctypes.c_int),
                ("replace_factor", ctypes.c_double), ("true_residual_stop", ctypes.c_int),
                ("restart", ctypes.c_int)]


class CgInfo(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("iters", ctypes.c_int), ("replacements", ctypes.c_int),
                ("pad", ctypes.c_int), ("res_norm", ctypes.c_double),
                ("true_res_norm", ctypes.c_double), ("b_norm", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "pad"}


class RequestTimes(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_double), ("admitted", ctypes.c_double), ("completed", ctypes.c_double)]


_lib = None
_vp = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
# double f(double x, double y, void *ctx)
FN2 = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_void_p)


def _declare(L):
    i = ctypes.c_int
    d = ctypes.c_double
    sz = ctypes.c_size_t
    i64 = ctypes.c_int64
    sig = {
        "sipg_strerror": (ctypes.c_char_p, [i]),
        "sipg_max_degree": (i, []),
        "sipg_max_fused_degree": (i, []),
        "sipg_basis_tables": (i, [i, _dp, _dp, _dp, _dp]),
        "sipg_device_count": (i, [_ip]),
        "sipg_set_device": (i, [i]),
        "sipg_device_synchronize": (i, []),
        "sipg_malloc": (i, [ctypes.POINTER(_vp), sz]),
        "sipg_free": (i, [_vp]),
        "sipg_memcpy_h2d": (i, [_vp, _vp, sz]),
        "sipg_memcpy_d2h": (i, [_vp, _vp, sz]),
        "sipg_memcpy_d2d": (i, [_vp, _vp, sz]),
        "sipg_memset_zero": (i, [_vp, sz]),
        "sipg_op_create": (i, [ctypes.POINTER(_vp), i, i, i, d, i, _vp]),
        "sipg_op_destroy": (i, [_vp]),
        "sipg_op_size": (i, [_vp, ctypes.POINTER(i64)]),
        "sipg_op_set_stream": (i, [_vp, _vp]),
        "sipg_op_tables": (i, [_vp, _dp, _dp, _dp, _dp]),
        "sipg_op_set_kernel": (i, [_vp, i]),
        "sipg_op_kernel": (i, [_vp, i, _ip]),
        "sipg_apply": (i, [_vp, _vp, _vp]),
        "sipg_dot": (i, [_vp, _vp, _vp, _dp]),
        "sipg_axpy": (i, [_vp, d, _vp, _vp]),
        "sipg_xpby": (i, [_vp, _vp, d, _vp]),
        "sipg_alpha_update": (i, [_vp, d, _vp, _vp, _vp, _vp]),
        "sipg_cg": (i, [_vp, _vp, _vp, d, i, i, _ip, _dp]),
        "sipg_true_relres": (i, [_vp, _vp, _vp, _dp]),
        "sipg_cg_ex": (i, [_vp, _vp, _vp, ctypes.POINTER(CgOpts), ctypes.POINTER(CgInfo)]),
        "sipg_cg_mixed": (i, [_vp, _vp, _vp, ctypes.POINTER(CgOpts), ctypes.POINTER(CgInfo)]),
        "sipg_pcg": (i, [_vp, _vp, _vp, ctypes.POINTER(CgOpts), ctypes.POINTER(CgInfo)]),
        "sipg_pcg_mixed": (i, [_vp, _vp, _vp, ctypes.POINTER(CgOpts), ctypes.POINTER(CgInfo)]),
        "sipg_prec_apply": (i, [_vp, _vp, _vp]),
        "sipg_prec2_apply": (i, [_vp, _vp, _vp]),
        "sipg_pcg2": (i, [_vp, _vp, _vp, ctypes.POINTER(CgOpts), ctypes.POINTER(CgInfo)]),
        "sipg_pcg2_mixed": (i, [_vp, _vp, _vp, ctypes.POINTER(CgOpts), ctypes.POINTER(CgInfo)]),
        "sipg_vec_dot": (i, [i, i64, _vp, _vp, _dp, _vp]),
        "sipg_vec_axpy": (i, [i, i64, d, _vp, _vp, _vp]),
        "sipg_vec_copy": (i, [i, i64, _vp, _vp, _vp]),
        "sipg_vec_xpby": (i, [i, i64, _vp, d, _vp, _vp]),
        "sipg_vec_alpha_update": (i, [i, i64, d, _vp, _vp, _vp, _vp, _vp]),
        "sipg_forcing_draw": (i, [ctypes.c_uint64, i, i, i, _ip, _dp, _ip, _ip]),
        "sipg_rhs_fourier": (i, [_vp, i, _dp, _ip, _ip, _vp]),
        "sipg_err_norms_fourier": (i, [_vp, _vp, i, _dp, _ip, _ip, _dp]),
        "sipg_write_result": (i, [_vp, _vp, ctypes.c_char_p]),
        "sipg_time_apply": (i, [_vp, _vp, _vp, i, _dp]),
        "sipg_upload": (i, [_vp, _dp, _vp]),
        "sipg_download": (i, [_vp, _vp, _dp]),
        "sipg_to_device_layout": (i, [_vp, _vp, _vp]),
        "sipg_to_reference_layout": (i, [_vp, _vp, _vp]),
        "sipg_batch_create": (i, [ctypes.POINTER(_vp), i, _ip, _ip, _ip, _dp, _vp]),
        "sipg_batch_destroy": (i, [_vp]),
        "sipg_batch_size": (i, [_vp, ctypes.POINTER(i64)]),
        "sipg_batch_offset": (i, [_vp, i, ctypes.POINTER(i64), ctypes.POINTER(i64)]),
        "sipg_batch_tune": (i, [_vp]),
        "sipg_batch_tuning": (i, [_vp, i, _ip, _ip, _dp]),
        "sipg_batch_set_shape": (i, [_vp, i, i, i]),
        "sipg_batch_set_graphs": (i, [_vp, i]),
        "sipg_batch_apply": (i, [_vp, _vp, _vp]),
        "sipg_batch_rhs_fourier": (i, [_vp, i, i, _dp, _ip, _ip, _vp]),
        "sipg_batch_upload": (i, [_vp, i, _dp, _vp]),
        "sipg_batch_download": (i, [_vp, i, _vp, _dp]),
        "sipg_batch_cg": (i, [_vp, _vp, _vp, ctypes.POINTER(CgOpts), ctypes.POINTER(CgInfo)]),
        "sipg_batch_pcg": (i, [_vp, _vp, _vp, ctypes.POINTER(CgOpts), ctypes.POINTER(CgInfo)]),
        "sipg_server_create": (i, [ctypes.POINTER(_vp), d, _vp]),
        "sipg_server_destroy": (i, [_vp]),
        "sipg_server_set_shape": (i, [_vp, i, i, i]),
        "sipg_server_set_method": (i, [_vp, i]),
        "sipg_server_submit": (i, [_vp, i, i, d, i, _dp, _ip, _ip, d, i, i, ctypes.POINTER(i64)]),
        "sipg_server_progress": (i, [_vp, _ip]),
        "sipg_server_take": (i, [_vp, ctypes.POINTER(i64), ctypes.POINTER(CgInfo), _dp, i64]),
        "sipg_server_stats": (i, [_vp, _ip, _ip, _ip, _dp]),
        "sipg_server_take_ex": (i, [_vp, ctypes.POINTER(i64), ctypes.POINTER(CgInfo), _dp, i64,
                                    ctypes.POINTER(RequestTimes)]),
        "sipg_server_cancel": (i, [_vp, i64]),
        "sipg_server_peek": (i, [_vp, ctypes.POINTER(i64), ctypes.POINTER(i64)]),
        "sipg_server_set_timeout": (i, [_vp, d]),
        "sipg_vec_dot_noalloc": (i, [i, i64, _vp, _vp, _dp, _vp]),
        "sipg_quad_points": (i, [_vp, _dp, _dp]),
        "sipg_rhs_nodal": (i, [_vp, _dp, _dp, _vp]),
        "sipg_rhs": (i, [_vp, FN2, FN2, _vp, _vp]),
        "sipg_err_norms_nodal": (i, [_vp, _vp, _dp, _dp, _dp, _dp]),
        "sipg_err_norms": (i, [_vp, _vp, FN2, FN2, FN2, _vp, _dp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    return L


def lib():
    """Load libsipg_hip.so from the package tree (raises if absent)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SipgLibraryMissing(
                f"{LIB_PATH} is not built; run `make -C {PKG_DIR}` or __graft_entry__.build()")
        _lib = _declare(ctypes.CDLL(LIB_PATH))
    return _lib


def strerror(status):
    return lib().sipg_strerror(status).decode()


def check(status, what=""):
    if status != OK:
        raise SipgError(status, what)
    return status


def max_degree():
    """Highest degree an operator accepts (63)."""
    return lib().sipg_max_degree()


def max_fused_degree():
    """Highest degree of the compiled fused operator kernels (16)."""
    return lib().sipg_max_fused_degree()


def device_count():
    c = ctypes.c_int(0)
    check(lib().sipg_device_count(ctypes.byref(c)), "device_count")
    return c.value


def synchronize():
    check(lib().sipg_device_synchronize(), "synchronize")


def _np_dtype(dtype):
    return np.float64 if dtype == F64 else np.float32


class DeviceBuffer:
    """A device allocation from the library allocator."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        self.ptr = _vp()
        check(lib().sipg_malloc(ctypes.byref(self.ptr), self.nbytes), "malloc")

    @property
    def value(self):
        return self.ptr.value

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes == self.nbytes
        check(lib().sipg_memcpy_h2d(self.ptr, arr.ctypes.data_as(_vp), self.nbytes), "h2d")

    def download(self, dtype, count=None):
        out = np.empty(self.nbytes // np.dtype(dtype).itemsize if count is None else count, dtype=dtype)
        check(lib().sipg_memcpy_d2h(out.ctypes.data_as(_vp), self.ptr, out.nbytes), "d2h")
        return out

    def zero(self):
        check(lib().sipg_memset_zero(self.ptr, self.nbytes), "memset")

    def free(self):
        if self.ptr is not None and self.ptr.value:
            lib().sipg_free(self.ptr)
        self.ptr = _vp()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Operator:
    """sipg_op: a dimx x dimy mesh, degree p, penalty pen, dtype."""

    def __init__(self, dimx, dimy, p, pen, dtype=F64, stream=None):
        self.dimx, self.dimy, self.p, self.pen, self.dtype = dimx, dimy, p, pen, dtype
        self.h = _vp()
        check(lib().sipg_op_create(ctypes.byref(self.h), dimx, dimy, p, float(pen), dtype,
                                   _vp(stream) if stream else None), "op_create")
        n = ctypes.c_int64()
        check(lib().sipg_op_size(self.h, ctypes.byref(n)), "op_size")
        self.n = n.value
        self.itemsize = 8 if dtype == F64 else 4

    def vector(self):
        return DeviceBuffer(self.n * self.itemsize)

    def upload(self, host, dev):
        host = np.ascontiguousarray(host, dtype=np.float64)
        assert host.size == self.n
        check(lib().sipg_upload(self.h, host.ctypes.data_as(_dp), dev.ptr), "upload")

    def download(self, dev):
        out = np.empty(self.n, dtype=np.float64)
        check(lib().sipg_download(self.h, dev.ptr, out.ctypes.data_as(_dp)), "download")
        return out

    def to_device(self, host):
        v = self.vector()
        self.upload(host, v)
        return v

    def to_device_layout(self, ref, dev):
        """Device array in the reference order -> element-major (async)."""
        return lib().sipg_to_device_layout(self.h, ref.ptr, dev.ptr)

    def to_reference_layout(self, dev, ref):
        """Element-major device array -> the reference order (async)."""
        return lib().sipg_to_reference_layout(self.h, dev.ptr, ref.ptr)

    def set_kernel(self, kernel):
        """KERNEL_AUTO, KERNEL_VALU or KERNEL_MFMA (sipg_op_set_kernel)."""
        check(lib().sipg_op_set_kernel(self.h, kernel), "op_set_kernel")

    def kernel(self, dtype=None):
        """The contraction kernel the op's applies (or, with dtype=F32 on an
        FP64 op, its mixed-precision inner cycles) use."""
        k = ctypes.c_int()
        check(lib().sipg_op_kernel(self.h, self.dtype if dtype is None else dtype, ctypes.byref(k)), "op_kernel")
        return k.value

    def apply(self, x, y):
        return lib().sipg_apply(self.h, x.ptr, y.ptr)

    def apply_host(self, xh):
        x = self.to_device(xh)
        y = self.vector()
        check(self.apply(x, y), "apply")
        out = self.download(y)
        x.free()
        y.free()
        return out

    def dot(self, a, b):
        r = ctypes.c_double()
        check(lib().sipg_dot(self.h, a.ptr, b.ptr, ctypes.byref(r)), "dot")
        return r.value

    def axpy(self, alpha, x, y):
        check(lib().sipg_axpy(self.h, alpha, x.ptr, y.ptr), "axpy")

    def xpby(self, r, beta, p):
        check(lib().sipg_xpby(self.h, r.ptr, beta, p.ptr), "xpby")

    def alpha_update(self, alpha, x, p, r, Ap):
        check(lib().sipg_alpha_update(self.h, alpha, x.ptr, p.ptr, r.ptr, Ap.ptr), "alpha_update")

    def cg(self, x, b, tol, tol_mode=TOL_REL, maxit=0):
        it = ctypes.c_int()
        res = ctypes.c_double()
        st = lib().sipg_cg(self.h, x.ptr, b.ptr, tol, tol_mode, maxit, ctypes.byref(it), ctypes.byref(res))
        return st, it.value, res.value

    def cg_ex(self, x, b, tol, tol_mode=TOL_REL, maxit=0, replace_factor=0.0, true_stop=True,
              restart=False):
        o = CgOpts(tol, tol_mode, maxit, replace_factor, 1 if true_stop else 0, int(restart))
        info = CgInfo()
        st = lib().sipg_cg_ex(self.h, x.ptr, b.ptr, ctypes.byref(o), ctypes.byref(info))
        return st, info

    def cg_mixed(self, x, b, tol, tol_mode=TOL_REL, maxit=0, replace_factor=0.0, restart=True):
        o = CgOpts(tol, tol_mode, maxit, replace_factor, 1, int(restart))
        info = CgInfo()
        st = lib().sipg_cg_mixed(self.h, x.ptr, b.ptr, ctypes.byref(o), ctypes.byref(info))
        return st, info

    def pcg(self, x, b, tol, tol_mode=TOL_REL, maxit=0, replace_factor=0.0, true_stop=True, restart=0):
        """Block-Jacobi preconditioned CG (sipg_pcg); options as cg_ex."""
        o = CgOpts(tol, tol_mode, maxit, replace_factor, 1 if true_stop else 0, int(restart))
        info = CgInfo()
        st = lib().sipg_pcg(self.h, x.ptr, b.ptr, ctypes.byref(o), ctypes.byref(info))
        return st, info

    def pcg2(self, x, b, tol, tol_mode=TOL_REL, maxit=0, replace_factor=0.0, true_stop=True, restart=0):
        """PCG with the two-level preconditioner (sipg_pcg2); options as pcg."""
        o = CgOpts(tol, tol_mode, maxit, replace_factor, 1 if true_stop else 0, int(restart))
        info = CgInfo()
        st = lib().sipg_pcg2(self.h, x.ptr, b.ptr, ctypes.byref(o), ctypes.byref(info))
        return st, info

    def pcg2_mixed(self, x, b, tol, tol_mode=TOL_REL, maxit=0, replace_factor=0.0, restart=True):
        """FP32 two-level PCG cycles inside FP64 refinement (sipg_pcg2_mixed)."""
        o = CgOpts(tol, tol_mode, maxit, replace_factor, 1, int(restart))
        info = CgInfo()
        st = lib().sipg_pcg2_mixed(self.h, x.ptr, b.ptr, ctypes.byref(o), ctypes.byref(info))
        return st, info

    def prec2_apply(self, r, z):
        """z = M2^-1 r with the two-level preconditioner of pcg2."""
        check(lib().sipg_prec2_apply(self.h, r.ptr, z.ptr), "prec2_apply")

    def pcg_mixed(self, x, b, tol, tol_mode=TOL_REL, maxit=0, replace_factor=0.0, restart=True):
        """FP32 block-Jacobi PCG cycles inside FP64 refinement (sipg_pcg_mixed)."""
        o = CgOpts(tol, tol_mode, maxit, replace_factor, 1, int(restart))
        info = CgInfo()
        st = lib().sipg_pcg_mixed(self.h, x.ptr, b.ptr, ctypes.byref(o), ctypes.byref(info))
        return st, info

    def prec_apply(self, r, z):
        """z = M^-1 r with the block-Jacobi preconditioner of pcg."""
        check(lib().sipg_prec_apply(self.h, r.ptr, z.ptr), "prec_apply")

    def true_relres(self, x, b):
        r = ctypes.c_double()
        check(lib().sipg_true_relres(self.h, x.ptr, b.ptr, ctypes.byref(r)), "true_relres")
        return r.value

    def rhs_fourier(self, modes, b):
        a, m, n = modes
        a = np.ascontiguousarray(a, dtype=np.float64)
        m = np.ascontiguousarray(m, dtype=np.int32)
        n = np.ascontiguousarray(n, dtype=np.int32)
        check(lib().sipg_rhs_fourier(self.h, len(a), a.ctypes.data_as(_dp), m.ctypes.data_as(_ip),
                                     n.ctypes.data_as(_ip), b.ptr), "rhs_fourier")

    def err_norms_fourier(self, modes, x):
        """Max nodal, L2 and H1 errors of x against the Fourier solution
        (sipg_sem_2d::err_norms), on the device."""
        a, m, n = modes
        a = np.ascontiguousarray(a, dtype=np.float64)
        m = np.ascontiguousarray(m, dtype=np.int32)
        n = np.ascontiguousarray(n, dtype=np.int32)
        out = np.zeros(3)
        check(lib().sipg_err_norms_fourier(self.h, x.ptr, len(a), a.ctypes.data_as(_dp), m.ctypes.data_as(_ip),
                                           n.ctypes.data_as(_ip), out.ctypes.data_as(_dp)), "err_norms_fourier")
        return {"max": out[0], "L2": out[1], "H1": out[2]}

    def quad_points(self):
        """Physical (x, y) of every GLL node, element-major like device
        vectors (sipg_quad_points)."""
        x, y = np.zeros(self.n), np.zeros(self.n)
        check(lib().sipg_quad_points(self.h, x.ctypes.data_as(_dp), y.ctypes.data_as(_dp)), "quad_points")
        return x, y

    def _nodal(self, v):
        """A function of (x, y) (vectorised over numpy arrays) or an array
        of nodal values, as a contiguous float64 array of the op's length."""
        if callable(v):
            x, y = self.quad_points()
            v = v(x, y)
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(v, dtype=np.float64), (self.n,)))
        return a

    def rhs(self, f, g=None, b=None):
        """Load vector of -lap u = f with Dirichlet data g (sipg_rhs_nodal;
        sipg_sem_2d::compute_rhs).  f, g: numpy-vectorised functions of
        (x, y) or nodal arrays; g None: g = 0.  Returns b (a new device
        vector unless one is given)."""
        fa = self._nodal(f)
        ga = None if g is None else self._nodal(g)
        if b is None:
            b = self.vector()
        check(lib().sipg_rhs_nodal(self.h, fa.ctypes.data_as(_dp),
                                   None if ga is None else ga.ctypes.data_as(_dp), b.ptr), "rhs_nodal")
        return b

    def rhs_fn(self, f, g=None, b=None):
        """As rhs with f and g as scalar Python functions called back from
        the library at each node (sipg_rhs)."""
        cf = FN2(lambda x, y, _c: f(x, y))
        cg = FN2(lambda x, y, _c: g(x, y)) if g is not None else FN2()
        if b is None:
            b = self.vector()
        check(lib().sipg_rhs(self.h, cf, cg, None, b.ptr), "rhs")
        return b

    def err_norms(self, x, u, ux, uy):
        """Max nodal, L2 and H1 errors of x against an exact solution u with
        derivatives ux, uy (numpy-vectorised functions or nodal arrays;
        sipg_err_norms_nodal)."""
        a, ax, ay = self._nodal(u), self._nodal(ux), self._nodal(uy)
        out = np.zeros(3)
        check(lib().sipg_err_norms_nodal(self.h, x.ptr, a.ctypes.data_as(_dp), ax.ctypes.data_as(_dp),
                                         ay.ctypes.data_as(_dp), out.ctypes.data_as(_dp)), "err_norms_nodal")
        return {"max": out[0], "L2": out[1], "H1": out[2]}

    def err_norms_fn(self, x, u, ux, uy):
        """As err_norms with scalar Python functions called back from the
        library (sipg_err_norms)."""
        cu, cx, cy = (FN2(lambda X, Y, _c, fn=fn: fn(X, Y)) for fn in (u, ux, uy))
        out = np.zeros(3)
        check(lib().sipg_err_norms(self.h, x.ptr, cu, cx, cy, None, out.ctypes.data_as(_dp)), "err_norms")
        return {"max": out[0], "L2": out[1], "H1": out[2]}

    def write_result(self, x, path=None):
        """x y u per node as text (sipg_sem_2d::print_result); None: stderr."""
        check(lib().sipg_write_result(self.h, x.ptr, None if path is None else os.fsencode(path)),
              "write_result")

    def time_apply(self, x, y, reps):
        ms = ctypes.c_double()
        check(lib().sipg_time_apply(self.h, x.ptr, y.ptr, reps, ctypes.byref(ms)), "time_apply")
        return ms.value

    def tables(self):
        n1 = self.p + 1
        z, w = np.zeros(n1), np.zeros(n1)
        D, K = np.zeros(n1 * n1), np.zeros(n1 * n1)
        check(lib().sipg_op_tables(self.h, z.ctypes.data_as(_dp), w.ctypes.data_as(_dp),
                                   D.ctypes.data_as(_dp), K.ctypes.data_as(_dp)), "tables")
        return z, w, D.reshape(n1, n1), K.reshape(n1, n1)

    def close(self):
        if self.h is not None and self.h.value:
            lib().sipg_op_destroy(self.h)
        self.h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def forcing_draw(seed, kmin=4, kmax=4, mmax=4):
    K = max(kmax, 1)
    a = np.zeros(K)
    m = np.zeros(K, dtype=np.int32)
    n = np.zeros(K, dtype=np.int32)
    nm = ctypes.c_int()
    check(lib().sipg_forcing_draw(ctypes.c_uint64(seed), kmin, kmax, mmax, ctypes.byref(nm),
                                  a.ctypes.data_as(_dp), m.ctypes.data_as(_ip), n.ctypes.data_as(_ip)),
          "forcing_draw")
    k = nm.value
    return a[:k].copy(), m[:k].copy(), n[:k].copy()


def basis_tables(p):
    """Host-only LGL nodes/weights, D[mode, node] and K1 for degree p."""
    n1 = p + 1
    z, w = np.zeros(n1), np.zeros(n1)
    D, K = np.zeros(n1 * n1), np.zeros(n1 * n1)
    check(lib().sipg_basis_tables(p, z.ctypes.data_as(_dp), w.ctypes.data_as(_dp),
                                  D.ctypes.data_as(_dp), K.ctypes.data_as(_dp)), "basis_tables")
    return z, w, D.reshape(n1, n1), K.reshape(n1, n1)


def pen_sigma(p):
    """Kernel penalty for sigma = 10 (p+1)^2 / h: pen = sigma/2 = 5 (p+1)^2."""
    return 5.0 * (p + 1) ** 2


class Batch:
    """Many independent FP64 problems solved together (sipg_batch_*).
    Instance i: a dims[i][0] x dims[i][1] mesh of degree degrees[i]."""

    def __init__(self, dimx, dimy, degrees, pens, stream=None):
        ni = len(degrees)
        self.ninst = ni
        self.dimx = np.ascontiguousarray(dimx, dtype=np.int32)
        self.dimy = np.ascontiguousarray(dimy, dtype=np.int32)
        self.p = np.ascontiguousarray(degrees, dtype=np.int32)
        self.pen = np.ascontiguousarray(pens, dtype=np.float64)
        self.h = _vp()
        check(lib().sipg_batch_create(ctypes.byref(self.h), ni, self.dimx.ctypes.data_as(_ip),
                                      self.dimy.ctypes.data_as(_ip), self.p.ctypes.data_as(_ip),
                                      self.pen.ctypes.data_as(_dp), _vp(stream) if stream else None),
              "batch_create")
        n = ctypes.c_int64()
        check(lib().sipg_batch_size(self.h, ctypes.byref(n)), "batch_size")
        self.n = n.value
        self.offsets = np.zeros(ni + 1, dtype=np.int64)
        for i in range(ni):
            o, k = ctypes.c_int64(), ctypes.c_int64()
            check(lib().sipg_batch_offset(self.h, i, ctypes.byref(o), ctypes.byref(k)), "batch_offset")
            self.offsets[i] = o.value
        self.offsets[ni] = self.n

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.sipg_batch_destroy(self.h)
            self.h = None

    def vector(self):
        return DeviceBuffer(self.n * 8)

    def tune(self):
        check(lib().sipg_batch_tune(self.h), "batch_tune")

    def tuning(self, p):
        bx, by, us = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        check(lib().sipg_batch_tuning(self.h, p, ctypes.byref(bx), ctypes.byref(by), ctypes.byref(us)),
              "batch_tuning")
        return bx.value, by.value, us.value

    def set_shape(self, p, bx, by):
        return lib().sipg_batch_set_shape(self.h, p, bx, by)

    def set_graphs(self, use_graphs):
        """True (default): replay captured CG iterations; False: launch them
        directly (sipg_batch_set_graphs)."""
        check(lib().sipg_batch_set_graphs(self.h, 1 if use_graphs else 0), "batch_set_graphs")

    def apply(self, x, y):
        return lib().sipg_batch_apply(self.h, x.ptr, y.ptr)

    def rhs_fourier(self, inst, modes, b):
        a, m, n = modes
        a = np.ascontiguousarray(a, dtype=np.float64)
        m = np.ascontiguousarray(m, dtype=np.int32)
        n = np.ascontiguousarray(n, dtype=np.int32)
        check(lib().sipg_batch_rhs_fourier(self.h, inst, len(a), a.ctypes.data_as(_dp), m.ctypes.data_as(_ip),
                                           n.ctypes.data_as(_ip), b.ptr), "batch_rhs_fourier")

    def upload(self, inst, host, dev):
        host = np.ascontiguousarray(host, dtype=np.float64)
        check(lib().sipg_batch_upload(self.h, inst, host.ctypes.data_as(_dp), dev.ptr), "batch_upload")

    def download(self, inst, dev):
        out = np.empty(int(self.offsets[inst + 1] - self.offsets[inst]))
        check(lib().sipg_batch_download(self.h, inst, dev.ptr, out.ctypes.data_as(_dp)), "batch_download")
        return out

    def cg(self, x, b, tol, tol_mode=TOL_REL, maxit=0, true_stop=True, replace_factor=0.0):
        o = CgOpts(tol, tol_mode, maxit, replace_factor, 1 if true_stop else 0, 0)
        infos = (CgInfo * self.ninst)()
        st = lib().sipg_batch_cg(self.h, x.ptr, b.ptr, ctypes.byref(o), infos)
        return st, list(infos)

    def pcg(self, x, b, tol, tol_mode=TOL_REL, maxit=0, true_stop=True, replace_factor=0.0):
        o = CgOpts(tol, tol_mode, maxit, replace_factor, 1 if true_stop else 0, 0)
        infos = (CgInfo * self.ninst)()
        st = lib().sipg_batch_pcg(self.h, x.ptr, b.ptr, ctypes.byref(o), infos)
        return st, list(infos)


class Server:
    """Continuous batching of independent FP64 solves (sipg_server_*)."""

    def __init__(self, hbm_budget_bytes, stream=None):
        self.h = _vp()
        check(lib().sipg_server_create(ctypes.byref(self.h), float(hbm_budget_bytes), _vp(stream) if stream else None),
              "server_create")

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.sipg_server_destroy(self.h)
            self.h = None

    def set_shape(self, p, bx, by):
        return lib().sipg_server_set_shape(self.h, p, bx, by)

    def set_method(self, method):
        """METHOD_CG or METHOD_PCG, while the server holds no request."""
        return lib().sipg_server_set_method(self.h, method)

    def submit(self, dim, p, pen, modes, tol, tol_mode=TOL_REL, maxit=0):
        a, m, n = modes
        a = np.ascontiguousarray(a, dtype=np.float64)
        m = np.ascontiguousarray(m, dtype=np.int32)
        n = np.ascontiguousarray(n, dtype=np.int32)
        rid = ctypes.c_int64()
        check(lib().sipg_server_submit(self.h, dim, p, pen, len(a), a.ctypes.data_as(_dp), m.ctypes.data_as(_ip),
                                       n.ctypes.data_as(_ip), tol, tol_mode, maxit, ctypes.byref(rid)),
              "server_submit")
        return rid.value

    def progress(self):
        nd = ctypes.c_int()
        check(lib().sipg_server_progress(self.h, ctypes.byref(nd)), "server_progress")
        return nd.value

    def take(self, n_solution=0):
        """(id, CgInfo, solution or None) of the oldest completed request, or None."""
        rid = ctypes.c_int64()
        info = CgInfo()
        x = np.empty(n_solution) if n_solution else None
        st = lib().sipg_server_take(self.h, ctypes.byref(rid), ctypes.byref(info),
                                    x.ctypes.data_as(_dp) if x is not None else None, n_solution)
        if st == ERR_NOT_FOUND:
            return None
        check(st, "server_take")
        return rid.value, info, x

    def peek(self):
        """(id, unknowns) of the oldest completed request, or None."""
        rid, n = ctypes.c_int64(), ctypes.c_int64()
        st = lib().sipg_server_peek(self.h, ctypes.byref(rid), ctypes.byref(n))
        if st == ERR_NOT_FOUND:
            return None
        check(st, "server_peek")
        return rid.value, n.value

    def take_ex(self, n_solution=0):
        """take() plus the request's times {submitted, admitted, completed}
        (seconds since the server's creation; admitted -1: never ran).
        n_solution = -1: a buffer of exactly the request's size."""
        if n_solution < 0:
            pk = self.peek()
            if pk is None:
                return None
            n_solution = pk[1]
        rid = ctypes.c_int64()
        info = CgInfo()
        t = RequestTimes()
        x = np.empty(n_solution) if n_solution else None
        st = lib().sipg_server_take_ex(self.h, ctypes.byref(rid), ctypes.byref(info),
                                       x.ctypes.data_as(_dp) if x is not None else None, n_solution, ctypes.byref(t))
        if st == ERR_NOT_FOUND:
            return None
        check(st, "server_take_ex")
        return rid.value, info, x, {"submitted": t.submitted, "admitted": t.admitted, "completed": t.completed}

    def cancel(self, rid):
        """Status of sipg_server_cancel (OK or ERR_NOT_FOUND)."""
        return lib().sipg_server_cancel(self.h, rid)

    def set_timeout(self, seconds):
        check(lib().sipg_server_set_timeout(self.h, float(seconds)), "server_set_timeout")

    def stats(self):
        q, r, d = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        b = ctypes.c_double()
        check(lib().sipg_server_stats(self.h, ctypes.byref(q), ctypes.byref(r), ctypes.byref(d), ctypes.byref(b)),
              "server_stats")
        return {"queued": q.value, "running": r.value, "done_held": d.value, "bytes_used": b.value}
