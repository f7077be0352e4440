"""The matrix-core operator apply (FP32, p = 15: 16 x 16 LGL nodes, one
v_mfma_f32_16x16x4f32 accumulator per element) against the oracle, and
against the vector-ALU kernel it stands in for.  FP32 bar as for the other
FP32 applies: max|y - y_oracle| <= 2e-6 (p+1) max|y_oracle|, the oracle
applied in FP64 to the same FP32-rounded input."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
P = 15
PEN = 5.0 * (P + 1) ** 2
TOL32 = 2e-6 * (P + 1)


def _x32(n, seed):
    return np.random.default_rng(seed).uniform(-1, 1, n).astype(np.float32).astype(np.float64)


@pytest.mark.parametrize("dx,dy", [(1, 1), (2, 1), (1, 3), (3, 2), (5, 3), (17, 9), (4, 33), (64, 64)])
@pytest.mark.parametrize("kernel", ["MFMA", "VALU"])
def test_p15_f32_apply_matches_oracle(sipg, oracle, dx, dy, kernel):
    op = sipg.Operator(dx, dy, P, PEN, sipg.F32)
    op.set_kernel(getattr(sipg, "KERNEL_" + kernel))
    assert op.kernel() == getattr(sipg, "KERNEL_" + kernel)
    x = _x32(op.n, 31 * dx + dy)
    y = op.apply_host(x)
    ref = oracle.apply(dx, dy, P, PEN, x)
    err = np.max(np.abs(y - ref)) / np.max(np.abs(ref))
    assert err <= TOL32, err


def test_p15_mfma_and_valu_agree_at_c3_size(sipg):
    # 128 x 128 elements, p = 15 (4,194,304 dof): the two kernels sum in
    # different orders; both are FP32 evaluations of the same operator
    op = sipg.Operator(128, 128, P, PEN, sipg.F32)
    x = _x32(op.n, 5)
    op.set_kernel(sipg.KERNEL_MFMA)
    ym = op.apply_host(x)
    op.set_kernel(sipg.KERNEL_VALU)
    yv = op.apply_host(x)
    scale = np.max(np.abs(yv))
    assert np.max(np.abs(ym - yv)) <= TOL32 * scale


def test_p15_mfma_apply_deterministic_and_symmetric(sipg):
    op = sipg.Operator(9, 6, P, PEN, sipg.F32)
    op.set_kernel(sipg.KERNEL_MFMA)
    x, z = _x32(op.n, 1), _x32(op.n, 2)
    ax = op.apply_host(x)
    assert np.array_equal(ax, op.apply_host(x))
    az = op.apply_host(z)
    # x.Az = z.Ax up to FP32 rounding of the two applies
    assert abs(z @ ax - x @ az) <= 1e-5 * np.sqrt((ax @ ax) * (z @ z))


def test_kernel_selection(sipg):
    op32 = sipg.Operator(4, 4, P, PEN, sipg.F32)
    assert op32.kernel() == sipg.KERNEL_MFMA              # the default picks the matrix cores here
    op64 = sipg.Operator(4, 4, P, PEN, sipg.F64)
    assert op64.kernel() == sipg.KERNEL_VALU              # no FP64 matrix-core kernel
    assert op64.kernel(sipg.F32) == sipg.KERNEL_MFMA      # its FP32 inner cycles
    op64.set_kernel(sipg.KERNEL_VALU)
    assert op64.kernel(sipg.F32) == sipg.KERNEL_VALU
    op8 = sipg.Operator(4, 4, 8, 405.0, sipg.F32)
    op8.set_kernel(sipg.KERNEL_MFMA)
    assert op8.kernel() == sipg.KERNEL_VALU               # p = 8 has no matrix-core kernel
    assert sipg.lib().sipg_op_set_kernel(op8.h, 7) == sipg.ERR_INVALID


@pytest.mark.parametrize("kernel", ["MFMA", "VALU"])
def test_p15_f32_cg_uses_the_fused_dot(sipg, oracle, kernel):
    # the apply with p.Ap drives the FP32 CG iteration: it must converge to
    # the FP32 level like the vector-ALU kernel does
    dim = 6
    op = sipg.Operator(dim, dim, P, PEN, sipg.F32)
    op.set_kernel(getattr(sipg, "KERNEL_" + kernel))
    b = op.to_device(oracle.rhs(dim, P, PEN, oracle.problem(0, sipg.forcing_draw(0, 4, 4, 4))))
    x = op.vector()
    x.zero()
    st, info = op.cg_ex(x, b, 1e-5, sipg.TOL_REL, maxit=20000, replace_factor=0.0, true_stop=False)
    assert st == sipg.OK, sipg.strerror(st)
    assert info.res_norm <= 1e-5 * info.b_norm
