"""GPU parity of the fused operator apply against the oracle (the
reference's volume + flux_term6a + flux_term6b restated).  FP64 bar:
max|y_gpu - y_oracle| <= 1e-12 * max|y_oracle|."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")
TOL64 = 1e-12


def _x(n, seed):
    return np.random.default_rng(seed).uniform(-1, 1, n)


@pytest.mark.parametrize("p", list(range(1, 17)))
# (50, 14) has workgroups whose halo ring lies inside the mesh on all sides
@pytest.mark.parametrize("dx,dy", [(5, 3), (17, 9), (1, 1), (16, 4), (33, 2), (50, 14)])
def test_apply_f64_matches_oracle(sipg, oracle, p, dx, dy):
    pen = 5.0 * (p + 1) ** 2
    op = sipg.Operator(dx, dy, p, pen, sipg.F64)
    x = _x(op.n, 100 * p + dx)
    y = op.apply_host(x)
    ref = oracle.apply(dx, dy, p, pen, x)
    err = np.max(np.abs(y - ref)) / np.max(np.abs(ref))
    assert err <= TOL64, err


@pytest.mark.parametrize("p", [2, 3, 8, 12, 15, 16])
@pytest.mark.parametrize("dx,dy", [(7, 5), (16, 8), (50, 14)])
def test_apply_f32_close_to_oracle(sipg, oracle, p, dx, dy):
    pen = 5.0 * (p + 1) ** 2
    op = sipg.Operator(dx, dy, p, pen, sipg.F32)
    x = _x(op.n, 7 * p + dy).astype(np.float32).astype(np.float64)
    y = op.apply_host(x)
    ref = oracle.apply(dx, dy, p, pen, x)
    err = np.max(np.abs(y - ref)) / np.max(np.abs(ref))
    # FP32 arithmetic: sums of O(p) terms of size O(pen)
    assert err <= 2e-6 * (p + 1), err


@pytest.mark.parametrize("name", ["d4p3", "d3p8", "d2p15", "r5x3p2"])
def test_apply_golden_vectors(sipg, name):
    g = np.load(os.path.join(GOLD, "apply_vectors.npz"))
    dx, dy, p, pen = g[name + "_shape"]
    op = sipg.Operator(int(dx), int(dy), int(p), float(pen), sipg.F64)
    y = op.apply_host(g[name + "_x"])
    ref = g[name + "_y"]
    assert np.max(np.abs(y - ref)) <= TOL64 * np.max(np.abs(ref))


def test_apply_full_size_c2(sipg, oracle):
    # configuration C2: 256 x 256 elements, p = 8, pen = 5 (p+1)^2
    dim, p = 256, 8
    pen = 5.0 * (p + 1) ** 2
    op = sipg.Operator(dim, dim, p, pen, sipg.F64)
    x = _x(op.n, 1)
    y = op.apply_host(x)
    ref = oracle.apply(dim, dim, p, pen, x)
    assert np.max(np.abs(y - ref)) <= TOL64 * np.max(np.abs(ref))


def test_apply_deterministic_and_linear(sipg):
    op = sipg.Operator(24, 20, 6, 245.0, sipg.F64)
    x1, x2 = _x(op.n, 1), _x(op.n, 2)
    y1 = op.apply_host(x1)
    assert np.array_equal(y1, op.apply_host(x1))
    y2 = op.apply_host(x2)
    y12 = op.apply_host(2.0 * x1 - 3.0 * x2)
    assert np.max(np.abs(y12 - (2 * y1 - 3 * y2))) <= 1e-12 * np.max(np.abs(y12))


def test_apply_symmetric(sipg):
    op = sipg.Operator(9, 7, 5, 180.0, sipg.F64)
    x, z = _x(op.n, 3), _x(op.n, 4)
    ax, az = op.apply_host(x), op.apply_host(z)
    assert abs(z @ ax - x @ az) <= 1e-12 * abs(z @ ax)


def test_apply_rejects_aliasing(sipg):
    op = sipg.Operator(4, 4, 3, 80.0, sipg.F64)
    v = op.to_device(_x(op.n, 5))
    assert op.apply(v, v) == sipg.ERR_ALIAS


def test_tables_match_oracle(sipg, oracle):
    for p in (1, 4, 8, 15, 16):
        op = sipg.Operator(2, 2, p, 1.0, sipg.F64)
        z, w, D, K = op.tables()
        assert np.max(np.abs(z - oracle.gll(p)[0])) < 1e-15 * (p + 1)
        assert np.max(np.abs(w - oracle.gll(p)[1])) < 1e-15 * (p + 1)
        Do = oracle.dphi(p)
        assert np.max(np.abs(D - Do)) < 1e-14 * np.max(np.abs(Do)) * (p + 1)
        Ko = oracle.k1(p)
        assert np.max(np.abs(K - Ko)) < 1e-14 * np.max(np.abs(Ko)) * (p + 1)


@pytest.mark.parametrize("dtype", ["F64", "F32"])
def test_layout_converters_round_trip(sipg, oracle, dtype):
    # a reference-order device array (a mode_vector's data) converted to the
    # element-major order must equal the host upload, and the operator applied
    # to it, converted back, must equal the oracle's reference-order result
    dim, p = 11, 5
    dt = getattr(sipg, dtype)
    op = sipg.Operator(dim, dim, p, 180.0, dt)
    npdt = np.float64 if dt == sipg.F64 else np.float32
    x = _x(op.n, 9).astype(npdt)
    ref_buf = sipg.DeviceBuffer(op.n * op.itemsize)
    ref_buf.upload(x)                                   # raw copy: reference order
    dev = op.vector()
    assert op.to_device_layout(ref_buf, dev) == 0
    assert np.array_equal(op.download(dev), x.astype(np.float64))
    y = op.vector()
    assert op.apply(dev, y) == 0
    back = sipg.DeviceBuffer(op.n * op.itemsize)
    assert op.to_reference_layout(y, back) == 0
    got = back.download(npdt).astype(np.float64)
    assert np.array_equal(got, op.download(y))
    ref = oracle.apply(dim, dim, p, 180.0, x.astype(np.float64))
    tol = TOL64 if dt == sipg.F64 else 2e-6 * (p + 1)
    assert np.max(np.abs(got - ref)) <= tol * np.max(np.abs(ref))
    assert op.to_device_layout(ref_buf, ref_buf) == sipg.ERR_ALIAS


# The vector-ALU kernel with whole elements per wave (p = 8): same
# per-line arithmetic as the workgroup-block kernel, so the two agree bit
# for bit; both against the oracle on meshes that cut blocks at every edge.
@pytest.mark.parametrize("p", [2, 4, 6, 8])
@pytest.mark.parametrize("dtype", ["F64", "F32"])
@pytest.mark.parametrize("dx,dy", [(1, 1), (5, 3), (7, 8), (8, 9), (15, 17), (33, 2), (50, 14), (128, 128)])
def test_wave_kernel_matches_oracle_and_block_kernel(sipg, oracle, p, dtype, dx, dy):
    pen = 5.0 * (p + 1) ** 2
    dt = getattr(sipg, dtype)
    op = sipg.Operator(dx, dy, p, pen, dt)
    x = _x(op.n, 13 * dx + dy)
    if dt == sipg.F32:
        x = x.astype(np.float32).astype(np.float64)
    op.set_kernel(sipg.KERNEL_WAVE)
    assert op.kernel() == sipg.KERNEL_WAVE
    yw = op.apply_host(x)
    op.set_kernel(sipg.KERNEL_VALU)
    yb = op.apply_host(x)
    assert np.array_equal(yw, yb)
    ref = oracle.apply(dx, dy, p, pen, x)
    tol = TOL64 if dt == sipg.F64 else 2e-6 * (p + 1)
    assert np.max(np.abs(yw - ref)) <= tol * np.max(np.abs(ref))


def test_wave_kernel_in_cg(sipg, oracle):
    # the fused p.Ap of the wave kernel drives CG like the block kernel's;
    # its block sums cover other element sets, so p.Ap is summed in another
    # order and the iterates agree to rounding, not bit for bit
    dim, p = 24, 8
    pen = 5.0 * (p + 1) ** 2
    op = sipg.Operator(dim, dim, p, pen, sipg.F64)
    b = op.to_device(oracle.rhs(dim, p, pen, oracle.problem(0, sipg.forcing_draw(0, 4, 4, 4))))
    res = {}
    for k in (sipg.KERNEL_WAVE, sipg.KERNEL_VALU):
        op.set_kernel(k)
        x = op.vector()
        x.zero()
        st, it, rn = op.cg(x, b, 1e-10, sipg.TOL_REL)
        assert st == sipg.OK
        res[k] = (it, op.download(x))
    assert abs(res[sipg.KERNEL_WAVE][0] - res[sipg.KERNEL_VALU][0]) <= 2
    xw, xb = res[sipg.KERNEL_WAVE][1], res[sipg.KERNEL_VALU][1]
    assert np.linalg.norm(xw - xb) <= 1e-9 * np.linalg.norm(xb)
