"""The general problem set-up through the library alone: node coordinates,
the load vector of any f with Dirichlet data g (sipg_sem_2d::compute_rhs,
sipg_sem_2d.hpp:232-305) and the error norms against any exact solution
(sipg_sem_2d::err_norms, sipg_sem_2d.hpp:311-370), checked against the
oracle's orc_rhs / orc_err_norms, and the thesis' Table 4.10 (solution E)
with its h-convergence rates reproduced without the oracle."""
import json
import math
import os

import numpy as np
import pytest

from analytic import SOLUTIONS

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def test_quad_points_follow_the_mesh_builder(sipg, oracle):
    dim, p = 7, 4
    op = sipg.Operator(dim, dim, p, 10.0, sipg.F64)
    x, y = op.quad_points()
    xe, ye = oracle.mesh_coords(dim)
    z, _ = oracle.gll(p)
    nodes = (z * .5 + .5) / dim
    n1 = p + 1
    X = (xe[:, None, None] + nodes[None, :, None] + 0 * nodes[None, None, :]).ravel()
    Y = (ye[:, None, None] + 0 * nodes[None, :, None] + nodes[None, None, :]).ravel()
    assert x.shape == (dim * dim * n1 * n1,)
    assert np.max(np.abs(x - X)) <= 4e-16 and np.max(np.abs(y - Y)) <= 4e-16


@pytest.mark.parametrize("kind,dim,p", [(1, 4, 2), (2, 5, 3), (3, 8, 4), (4, 3, 6), (5, 16, 2), (5, 1, 3),
                                        (2, 9, 8), (1, 6, 15)])
def test_rhs_matches_the_oracle(sipg, oracle, kind, dim, p):
    u, f, _, _ = SOLUTIONS[kind]
    pen = 100.0 * p * p
    op = sipg.Operator(dim, dim, p, pen, sipg.F64)
    b = op.rhs(f, u)
    want = oracle.rhs(dim, p, pen, oracle.problem(kind))
    got = op.download(b)
    scale = np.abs(want).max()
    assert np.max(np.abs(got - want)) <= 1e-13 * scale, np.max(np.abs(got - want)) / scale


def test_rhs_homogeneous_data_is_the_volume_term(sipg, oracle):
    # g = 0: only (w h/2)(w h/2) f, which the Fourier forcing's load vector
    # (zero on the boundary) also is, up to the rounding of sin(m pi) at x = 1
    modes = sipg.forcing_draw(3, 4, 4, 4)
    a, m, n = modes
    f = lambda x, y: sum(ak * math.pi ** 2 * (mk * mk + nk * nk) * np.sin(mk * math.pi * x) * np.sin(nk * math.pi * y)
                         for ak, mk, nk in zip(a, m, n))
    op = sipg.Operator(12, 12, 5, sipg.pen_sigma(5), sipg.F64)
    b1 = op.download(op.rhs(f))
    b2 = op.vector()
    op.rhs_fourier(modes, b2)
    b2 = op.download(b2)
    assert np.max(np.abs(b1 - b2)) <= 1e-13 * np.abs(b2).max()


def test_callback_forms_match_the_nodal_forms(sipg):
    kind, dim, p = 1, 5, 3
    u, f, ux, uy = SOLUTIONS[kind]
    op = sipg.Operator(dim, dim, p, 90.0, sipg.F64)
    scal = lambda fn: (lambda x, y: float(fn(np.float64(x), np.float64(y))))
    # (numpy's vector and scalar exp may differ in the last bit)
    b1 = op.download(op.rhs(f, u))
    b2 = op.download(op.rhs_fn(scal(f), scal(u)))
    assert np.max(np.abs(b1 - b2)) <= 1e-15 * np.abs(b1).max()
    x = op.vector()
    x.zero()
    op.cg(x, op.to_device(b1), 1e-13, sipg.TOL_ABS)
    n1 = op.err_norms(x, u, ux, uy)
    n2 = op.err_norms_fn(x, scal(u), scal(ux), scal(uy))
    for k in ("max", "L2", "H1"):
        assert abs(n1[k] - n2[k]) <= 1e-13 * n1[k], k


@pytest.mark.parametrize("kind,dim,p,dtype", [(1, 8, 3, "F64"), (2, 6, 5, "F64"), (5, 16, 2, "F64"),
                                              (3, 4, 4, "F32"), (5, 12, 8, "F64")])
def test_err_norms_match_the_oracle(sipg, oracle, kind, dim, p, dtype):
    u, f, ux, uy = SOLUTIONS[kind]
    pen = 100.0 * p * p
    dt = getattr(sipg, dtype)
    op = sipg.Operator(dim, dim, p, pen, dt)
    b = op.rhs(f, u)
    x = op.vector()
    x.zero()
    op.cg(x, b, 1e-12 if dt == sipg.F64 else 1e-4, sipg.TOL_ABS if dt == sipg.F64 else sipg.TOL_REL)
    ours = op.err_norms(x, u, ux, uy)
    want = oracle.err_norms(dim, p, op.download(x), oracle.problem(kind))
    # the exact solution is sampled by numpy here and by libm in the oracle
    # (last-bit differences), and the derivative u_x = 2 dim sum_i u_i D_i
    # sums terms of size 2 dim |D| |u| (~1e3 for E at p = 8) in another
    # order: absolute differences up to ~1e-13, hence the absolute floor
    for k in ("max", "L2", "H1"):
        assert abs(ours[k] - want[k]) <= 1e-12 * want[k] + 1e-12, (k, ours[k], want[k])


def _thesis_rows(sol):
    with open(os.path.join(GOLD, "thesis_tables.json")) as f:
        return json.load(f)["solutions"][sol]["rows"]


@pytest.mark.parametrize("p", [2, 3, 4])
def test_thesis_table_4_10_and_rates_through_the_library(sipg, p):
    # solution E, penalty 100 p^2, the thesis' absolute tolerance 1e-15:
    # RHS, solve and norms all by the library, errors within the 6 printed
    # digits, and the h-rates: L2 at least p + 1 (the nodal L2 of GLL
    # quadrature shows p + 2) and H1 -> p on the finest pair
    u, f, ux, uy = SOLUTIONS[5]
    rows = [r for r in _thesis_rows("E") if r["p"] == p]
    got = []
    for r in rows:
        dim = r["dim"]
        op = sipg.Operator(dim, dim, p, 100.0 * p * p, sipg.F64)
        b = op.rhs(f, u)
        x = op.vector()
        x.zero()
        st, info = op.pcg2(x, b, 1e-15, sipg.TOL_ABS, true_stop=False)
        assert st in (sipg.OK, sipg.NOT_CONVERGED), sipg.strerror(st)
        nrm = op.err_norms(x, u, ux, uy)
        for key, ours in (("H1", nrm["H1"]), ("L2", nrm["L2"]), ("Linf", nrm["max"])):
            assert abs(ours - r[key]) <= 2e-5 * r[key], (dim, key, ours, r[key])
        got.append(nrm)
    l2 = [math.log2(a["L2"] / b["L2"]) for a, b in zip(got, got[1:])]
    h1 = [math.log2(a["H1"] / b["H1"]) for a, b in zip(got, got[1:])]
    assert l2[-1] >= p + 1 - 0.05, l2
    assert abs(h1[-1] - p) <= 0.05, h1
