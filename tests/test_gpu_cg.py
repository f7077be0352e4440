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
        assert np.max(np.abs(got - ref)) <= 1e-13 * np.max(np.abs(ref))


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
