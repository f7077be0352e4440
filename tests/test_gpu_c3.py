"""Configuration C3: 128 x 128 elements of degree p = 15 (16 x 16 LGL nodes,
4,194,304 dof), the FP32 inner two-level PCG cycles on the matrix cores
inside FP64 refinement (sipg_pcg2_mixed) to a 1e-10 true relative residual;
seeded Fourier forcing with 1-8 modes, m, n <= 12 (seed 0).

On a reduced mesh (16 x 16, same degree, same forcing) the solution is
compared with the oracle's FP64 CG solve (the reference's
conjugate_gradient restated, run to 1e-13); at full size, with the FP64
two-level PCG solve and with the exact Fourier solution.  The measured
differences are printed."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
P = 15
TOL = 1e-10


def _problem(sipg, dim):
    op = sipg.Operator(dim, dim, P, sipg.pen_sigma(P), sipg.F64)
    modes = sipg.forcing_draw(0, 1, 8, 12)
    b = op.vector()
    op.rhs_fourier(modes, b)
    return op, modes, b


def _solve(sipg, op, b, method):
    x = op.vector()
    x.zero()
    if method == "mixed":
        st, info = op.pcg2_mixed(x, b, TOL, sipg.TOL_REL, maxit=60000, replace_factor=0.1, restart=1)
    else:
        st, info = op.pcg2(x, b, TOL, sipg.TOL_REL, maxit=30000, replace_factor=0.1, true_stop=True)
    assert st == sipg.OK, sipg.strerror(st)
    rel = op.true_relres(x, b)
    assert rel <= TOL, rel
    return x, info


@pytest.mark.parametrize("method,kernel", [("mixed", "MFMA"), ("mixed", "VALU"), ("fp64", "AUTO")])
def test_c3_reduced_mesh_matches_the_oracle_solve(sipg, oracle, method, kernel):
    dim = 16
    op, modes, b = _problem(sipg, dim)
    op.set_kernel(getattr(sipg, "KERNEL_" + kernel))
    if method == "mixed":
        assert op.kernel(sipg.F32) == getattr(sipg, "KERNEL_" + ("VALU" if kernel == "VALU" else "MFMA"))
    x, info = _solve(sipg, op, b, method)
    bh = op.download(b)
    xo, _, _ = oracle.cg(dim, dim, P, op.pen, bh, tol=1e-13, tol_mode=1)
    xh = op.download(x)
    l2 = np.linalg.norm(xh - xo) / np.linalg.norm(xo)
    mx = np.max(np.abs(xh - xo)) / np.max(np.abs(xo))
    print(f"C3 reduced ({dim}x{dim}, p={P}) {method}/{kernel}: {info.iters} iterations, "
          f"vs oracle: rel L2 {l2:.2e}, rel max {mx:.2e}")
    assert l2 <= 1e-8 and mx <= 1e-8, (l2, mx)


def test_c3_full_size_mixed_matches_fp64(sipg):
    op, modes, b = _problem(sipg, 128)
    assert op.kernel(sipg.F32) == sipg.KERNEL_MFMA
    x64, i64 = _solve(sipg, op, b, "fp64")
    xm, im = _solve(sipg, op, b, "mixed")
    h64, hm = op.download(x64), op.download(xm)
    l2 = np.linalg.norm(hm - h64) / np.linalg.norm(h64)
    mx = np.max(np.abs(hm - h64)) / np.max(np.abs(h64))
    e64, em = op.err_norms_fourier(modes, x64), op.err_norms_fourier(modes, xm)
    print(f"C3 128x128 p=15: FP64 {i64.iters} iterations, mixed {im.iters} FP32 iterations in "
          f"{im.replacements} cycles; mixed vs FP64: rel L2 {l2:.2e}, rel max {mx:.2e}; "
          f"error vs exact: FP64 max {e64['max']:.2e}, mixed max {em['max']:.2e}")
    assert l2 <= 1e-9 and mx <= 1e-9, (l2, mx)
    # the discretisation error at p = 15 is far below the solver tolerance's effect
    assert em["max"] <= 1e-9 and e64["max"] <= 1e-9
