"""GPU CG: vector kernels, the device-resident solve against the dense
LAPACK oracle (C1), the thesis error tables through the GPU solve, and the
forced failure paths."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def test_vector_kernels(sipg):
    op = sipg.Operator(6, 5, 4, 125.0, sipg.F64)
    rng = np.random.default_rng(0)
    a, b, c, d = (rng.uniform(-1, 1, op.n) for _ in range(4))
    A, B, C, Dv = (op.to_device(v) for v in (a, b, c, d))
    assert abs(op.dot(A, B) - a @ b) <= 1e-13 * np.abs(a * b).sum()
    op.axpy(0.5, A, B)
    assert np.array_equal(op.download(B), b + 0.5 * a)
    op.xpby(A, -2.0, C)            # C = A - 2 C
    assert np.array_equal(op.download(C), a + -2.0 * c)
    x, p, r, ap = (rng.uniform(-1, 1, op.n) for _ in range(4))
    X, P, R, AP = (op.to_device(v) for v in (x, p, r, ap))
    op.alpha_update(0.25, X, P, R, AP)
    assert np.array_equal(op.download(X), x + 0.25 * p)
    assert np.array_equal(op.download(R), r - 0.25 * ap)


def test_dot_deterministic(sipg):
    op = sipg.Operator(64, 64, 7, 320.0, sipg.F64)
    a = np.random.default_rng(1).uniform(-1, 1, op.n)
    A = op.to_device(a)
    d1 = op.dot(A, A)
    for _ in range(3):
        assert op.dot(A, A) == d1


def test_rhs_fourier_matches_oracle(sipg, oracle):
    modes = sipg.forcing_draw(0, 4, 4, 4)
    for dim, p in ((8, 3), (13, 6)):
        op = sipg.Operator(dim, dim, p, sipg.pen_sigma(p), sipg.F64)
        b = op.vector()
        op.rhs_fourier(modes, b)
        got = op.download(b)
        ref = oracle.rhs(dim, p, sipg.pen_sigma(p), oracle.problem(0, modes))
        # The reference evaluates the Dirichlet data g = u_ex at mesh
        # coordinates accumulated by repeated addition
        # (build_square_mesh.hpp:256-284); for h = 1/13 the bottom edge sits at
        # y = 2.8e-16, so its g = 0 boundary terms come out O(1e-13) instead of
        # 0.  The device rhs uses g = 0 exactly.
        assert np.max(np.abs(got - ref)) <= 1e-11 * np.max(np.abs(ref))


def test_c1_solve_matches_dense_lapack(sipg):
    g = np.load(os.path.join(GOLD, "c1_solution.npz"))
    op = sipg.Operator(8, 8, 3, 80.0, sipg.F64)
    b = op.to_device(g["rhs"])
    x = op.vector()
    x.zero()
    st, it, res = op.cg(x, b, 1e-10, sipg.TOL_REL)
    assert st == sipg.OK, sipg.strerror(st)
    u = op.download(x)
    assert np.linalg.norm(u - g["u"]) / np.linalg.norm(g["u"]) <= 1e-8
    assert op.true_relres(x, b) <= 1e-10 * 1.5
    assert res <= 1e-10 * np.linalg.norm(g["rhs"])


def test_cg_iterations_track_oracle(sipg, oracle):
    modes = sipg.forcing_draw(0, 4, 4, 4)
    dim, p = 16, 4
    pen = sipg.pen_sigma(p)
    bh = oracle.rhs(dim, p, pen, oracle.problem(0, modes))
    xo, ito, rro = oracle.cg(dim, dim, p, pen, bh, tol=1e-10, tol_mode=1)
    op = sipg.Operator(dim, dim, p, pen, sipg.F64)
    b = op.to_device(bh)
    x = op.vector()
    x.zero()
    st, it, res = op.cg(x, b, 1e-10, sipg.TOL_REL)
    assert st == sipg.OK
    assert abs(it - ito) <= max(2, 0.02 * ito), (it, ito)
    assert np.linalg.norm(op.download(x) - xo) <= 1e-8 * np.linalg.norm(xo)


# p = 14 and 16 in FP64 use 8x4-element workgroups of 480 and 544 threads:
# the fused p.Ap sum then ends in a partial wave
@pytest.mark.parametrize("p", [13, 14, 16])
def test_cg_partial_wave_degrees_track_oracle(sipg, oracle, p):
    modes = sipg.forcing_draw(0, 4, 4, 4)
    dim = 6
    pen = sipg.pen_sigma(p)
    bh = oracle.rhs(dim, p, pen, oracle.problem(0, modes))
    xo, ito, _ = oracle.cg(dim, dim, p, pen, bh, tol=1e-10, tol_mode=1)
    op = sipg.Operator(dim, dim, p, pen, sipg.F64)
    b = op.to_device(bh)
    x = op.vector()
    x.zero()
    st, it, _ = op.cg(x, b, 1e-10, sipg.TOL_REL)
    assert st == sipg.OK
    assert abs(it - ito) <= max(2, 0.02 * ito), (it, ito)
    assert np.linalg.norm(op.download(x) - xo) <= 1e-8 * np.linalg.norm(xo)


def test_cg_absolute_tolerance_semantics(sipg, oracle):
    dim, p, pen = 8, 2, 45.0
    bh = oracle.rhs(dim, p, pen, oracle.problem(5))
    op = sipg.Operator(dim, dim, p, pen, sipg.F64)
    b = op.to_device(bh)
    x = op.vector()
    x.zero()
    st, it, res = op.cg(x, b, 1e-6, sipg.TOL_ABS)
    assert st == sipg.OK and res <= 1e-6
    xo, ito, rro = oracle.cg(dim, dim, p, pen, bh, tol=1e-6, tol_mode=0)
    assert abs(it - ito) <= 2


@pytest.mark.parametrize("sol,dim,p", [("E", 16, 2), ("E", 8, 3), ("E", 16, 4), ("A", 8, 3)])
def test_thesis_tables_through_gpu_solve(sipg, oracle, sol, dim, p):
    with open(os.path.join(GOLD, "thesis_tables.json")) as f:
        th = json.load(f)["solutions"][sol]
    row = [r for r in th["rows"] if r["dim"] == dim and r["p"] == p][0]
    pen = 100.0 * p * p
    prob = oracle.problem(th["kind"])
    bh = oracle.rhs(dim, p, pen, prob)          # host rhs, like sipg_sem_2d::compute_rhs
    op = sipg.Operator(dim, dim, p, pen, sipg.F64)
    b = op.to_device(bh)
    x = op.vector()
    x.zero()
    st, it, res = op.cg(x, b, 1e-15, sipg.TOL_ABS)
    assert st in (sipg.OK, sipg.NOT_CONVERGED)
    nrm = oracle.err_norms(dim, p, op.download(x), prob)
    for key, ours in (("H1", nrm["H1"]), ("L2", nrm["L2"]), ("Linf", nrm["max"])):
        assert abs(ours - row[key]) <= 2e-5 * row[key], (key, ours, row[key])


def test_not_converged_reports_cap(sipg):
    op = sipg.Operator(8, 8, 3, 80.0, sipg.F64)
    b = op.to_device(np.random.default_rng(2).uniform(-1, 1, op.n))
    x = op.vector()
    x.zero()
    st, it, res = op.cg(x, b, 1e-12, sipg.TOL_REL, maxit=5)
    assert st == sipg.NOT_CONVERGED and it == 5


def test_nonfinite_rhs_detected(sipg):
    op = sipg.Operator(4, 4, 2, 45.0, sipg.F64)
    bh = np.ones(op.n)
    bh[7] = np.nan
    b = op.to_device(bh)
    x = op.vector()
    x.zero()
    st, it, res = op.cg(x, b, 1e-10, sipg.TOL_REL)
    assert st == sipg.ERR_NONFINITE and it == 0


def test_breakdown_on_indefinite_operator(sipg):
    # a negative penalty makes A indefinite (below coercivity); the first
    # search direction of a jump-heavy rhs has p.Ap < 0
    op = sipg.Operator(4, 4, 2, -500.0, sipg.F64)
    bh = np.random.default_rng(4).uniform(-1, 1, op.n)
    b = op.to_device(bh)
    x = op.vector()
    x.zero()
    st, it, res = op.cg(x, b, 1e-10, sipg.TOL_REL, maxit=200)
    assert st == sipg.ERR_BREAKDOWN


def test_cg_f32_converges_to_f32_level(sipg, oracle):
    modes = sipg.forcing_draw(0, 4, 4, 4)
    dim, p = 16, 5
    pen = sipg.pen_sigma(p)
    bh = oracle.rhs(dim, p, pen, oracle.problem(0, modes))
    op = sipg.Operator(dim, dim, p, pen, sipg.F32)
    b = op.to_device(bh)
    x = op.vector()
    x.zero()
    st, it, res = op.cg(x, b, 1e-5, sipg.TOL_REL)
    assert st == sipg.OK
    xo, _, _ = oracle.cg(dim, dim, p, pen, bh, tol=1e-12, tol_mode=1)
    assert np.linalg.norm(op.download(x) - xo) <= 1e-3 * np.linalg.norm(xo)


def test_cg_ex_reaches_true_relative_residual(sipg, oracle):
    modes = sipg.forcing_draw(0, 4, 4, 4)
    dim, p = 32, 6
    pen = sipg.pen_sigma(p)
    op = sipg.Operator(dim, dim, p, pen, sipg.F64)
    b = op.vector()
    op.rhs_fourier(modes, b)
    x = op.vector()
    x.zero()
    st, info = op.cg_ex(x, b, 1e-10, sipg.TOL_REL, replace_factor=0.1, true_stop=True)
    assert st == sipg.OK, sipg.strerror(st)
    assert info.replacements >= 2
    rel = op.true_relres(x, b)
    assert rel <= 1e-10, rel
    assert info.true_res_norm <= 1e-10 * info.b_norm * 1.0000001


def test_cg_ex_without_replacement_equals_cg(sipg):
    op = sipg.Operator(12, 12, 4, 125.0, sipg.F64)
    bh = np.random.default_rng(9).uniform(-1, 1, op.n)
    b = op.to_device(bh)
    x1, x2 = op.vector(), op.vector()
    x1.zero()
    x2.zero()
    st1, it1, res1 = op.cg(x1, b, 1e-9, sipg.TOL_REL)
    st2, info = op.cg_ex(x2, b, 1e-9, sipg.TOL_REL, replace_factor=0.0, true_stop=False)
    assert st1 == st2 == sipg.OK and it1 == info.iters
    assert np.array_equal(op.download(x1), op.download(x2))


def test_mixed_precision_matches_fp64_oracle(sipg, oracle):
    modes = sipg.forcing_draw(0, 4, 4, 4)
    dim, p = 24, 7
    pen = sipg.pen_sigma(p)
    bh = oracle.rhs(dim, p, pen, oracle.problem(0, modes))
    op = sipg.Operator(dim, dim, p, pen, sipg.F64)
    b = op.to_device(bh)
    x = op.vector()
    x.zero()
    st, info = op.cg_mixed(x, b, 1e-10, sipg.TOL_REL)
    assert st == sipg.OK, sipg.strerror(st)
    rel = op.true_relres(x, b)
    assert rel <= 1e-10, rel
    # error against the FP64 oracle solve at the same tolerance
    xo, ito, _ = oracle.cg(dim, dim, p, pen, bh, tol=1e-12, tol_mode=1)
    err = np.linalg.norm(op.download(x) - xo) / np.linalg.norm(xo)
    assert err <= 1e-8, err
    assert info.replacements >= 2


def test_mixed_precision_keep_direction_variant_finishes(sipg, oracle):
    # the reliable-update variant (search direction kept across true
    # residuals) on a small, well-conditioned case
    modes = sipg.forcing_draw(0, 4, 4, 4)
    dim, p = 8, 3
    pen = sipg.pen_sigma(p)
    bh = oracle.rhs(dim, p, pen, oracle.problem(0, modes))
    op = sipg.Operator(dim, dim, p, pen, sipg.F64)
    b = op.to_device(bh)
    x = op.vector()
    x.zero()
    st, info = op.cg_mixed(x, b, 1e-10, sipg.TOL_REL, replace_factor=0.1, restart=False)
    assert st == sipg.OK, sipg.strerror(st)
    assert op.true_relres(x, b) <= 1e-10


def test_operator_free_vector_kernels(sipg):
    import ctypes
    n = 10007
    a = np.random.default_rng(5).uniform(-1, 1, n)
    b = np.random.default_rng(6).uniform(-1, 1, n)
    A, B = sipg.DeviceBuffer(8 * n), sipg.DeviceBuffer(8 * n)
    A.upload(a)
    B.upload(b)
    r = ctypes.c_double()
    sipg.check(sipg.lib().sipg_vec_dot(sipg.F64, n, A.ptr, B.ptr, ctypes.byref(r), None), "vec_dot")
    assert abs(r.value - a @ b) <= 1e-13 * np.abs(a * b).sum()
    sipg.check(sipg.lib().sipg_vec_axpy(sipg.F64, n, 2.0, A.ptr, B.ptr, None), "vec_axpy")
    sipg.synchronize()
    assert np.array_equal(B.download(np.float64), b + 2.0 * a)
    C = sipg.DeviceBuffer(8 * n)
    sipg.check(sipg.lib().sipg_vec_copy(sipg.F64, n, A.ptr, C.ptr, None), "vec_copy")
    sipg.synchronize()
    assert np.array_equal(C.download(np.float64), a)
    assert sipg.lib().sipg_vec_copy(sipg.F64, n, A.ptr, A.ptr, None) == sipg.ERR_ALIAS


@pytest.mark.parametrize("dim,p", [(8, 3), (13, 6), (16, 2)])
def test_err_norms_match_oracle(sipg, oracle, dim, p):
    # sipg_err_norms_fourier against the oracle's err_norms (itself pinned by
    # the thesis error tables): on a random vector (errors of order one, so
    # the comparison is tight) and on a CG solution (the discretisation error)
    modes = sipg.forcing_draw(3, 4, 4, 4)
    prob = oracle.problem(0, modes)
    op = sipg.Operator(dim, dim, p, sipg.pen_sigma(p), sipg.F64)
    xr = np.random.default_rng(dim * p).uniform(-1, 1, op.n)
    x = op.to_device(xr)
    got = op.err_norms_fourier(modes, x)
    ref = oracle.err_norms(dim, p, xr, prob)
    for k in ("max", "L2", "H1"):
        assert abs(got[k] - ref[k]) <= 1e-12 * ref[k], (k, got[k], ref[k])
    b = op.vector()
    op.rhs_fourier(modes, b)
    x.zero()
    st, it, res = op.cg(x, b, 1e-13, sipg.TOL_REL)
    got = op.err_norms_fourier(modes, x)
    ref = oracle.err_norms(dim, p, op.download(x), prob)
    # the oracle places nodes on the reference's mesh coordinates, summed
    # edge by edge (O(1e-16) off the exact ones): 1e-5 of an error of 1e-9
    for k in ("max", "L2", "H1"):
        assert abs(got[k] - ref[k]) <= 1e-5 * ref[k], (k, got[k], ref[k])
    # convergence: a finer mesh of the same degree has a smaller error
    op2 = sipg.Operator(2 * dim, 2 * dim, p, sipg.pen_sigma(p), sipg.F64)
    b2, x2 = op2.vector(), op2.vector()
    op2.rhs_fourier(modes, b2)
    x2.zero()
    op2.cg(x2, b2, 1e-13, sipg.TOL_REL)
    fine = op2.err_norms_fourier(modes, x2)
    assert fine["L2"] < 0.25 * got["L2"] and fine["H1"] < 0.5 * got["H1"], (fine, got)


def test_err_norms_of_the_interpolant(sipg):
    # the nodal interpolant of u_ex has zero nodal error (max and L2 are
    # nodal quadratures), and its H1 error is the derivative error of the
    # degree-p interpolant, which falls with the mesh size.  The nodes are
    # placed by the reference's convention (element e = dim*row + col,
    # row 0 at the top: x = (col + (z_i+1)/2)/dim, y = 1 - (row + 1 - (z_j+1)/2)/dim)
    modes = sipg.forcing_draw(5, 4, 4, 4)
    a, m, n = modes
    h1 = []
    p = 5
    z, _, _, _ = sipg.basis_tables(p)
    for dim, dtype in ((6, sipg.F64), (12, sipg.F64), (6, sipg.F32)):
        e = np.arange(dim * dim)
        col, row = e % dim, e // dim
        X = (col[None, None, :] + (z[:, None, None] + 1) / 2) / dim            # [i1, 1, e]
        Y = 1 - (row[None, None, :] + 1 - (z[None, :, None] + 1) / 2) / dim    # [1, i2, e]
        u = sum(a[k] * np.sin(m[k] * np.pi * X) * np.sin(n[k] * np.pi * Y) for k in range(len(a)))
        op = sipg.Operator(dim, dim, p, sipg.pen_sigma(p), dtype)
        x = op.to_device(u.reshape(-1))                  # reference order (i1, i2, e)
        got = op.err_norms_fourier(modes, x)
        tol = 1e-14 if dtype == sipg.F64 else 1e-6       # FP32 rounding of the uploaded u
        assert got["max"] <= tol and got["L2"] <= tol, (dim, dtype, got)
        if dtype == sipg.F64:
            h1.append(got["H1"])
    assert 0 < h1[1] < 0.5 ** 4 * h1[0], h1


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_write_result_matches_oracle(sipg, oracle, tmp_path, dtype):
    # sipg_write_result is print_result's text: byte for byte the oracle's
    # file for the same values (the oracle's format is pinned on the CPU)
    dim, p = 13, 4
    op = sipg.Operator(dim, dim, p, sipg.pen_sigma(p), sipg.F64 if dtype == "f64" else sipg.F32)
    b, x = op.vector(), op.vector()
    op.rhs_fourier(sipg.forcing_draw(2, 4, 4, 4), b)
    x.zero()
    op.cg(x, b, 1e-6 if dtype == "f32" else 1e-12, sipg.TOL_REL)
    got, want = tmp_path / "gpu.txt", tmp_path / "oracle.txt"
    op.write_result(x, str(got))
    oracle.print_result(dim, p, op.download(x), str(want))
    assert got.read_bytes() == want.read_bytes()
