"""Batched solves (sipg_batch_*) against the single-problem operator and
solver, which the other GPU tests pin to the oracle.  The batched kernels run
the same per-block arithmetic, so the operator must agree bit for bit for
every block shape; the batched CG must reach the same true relative
residual, with per-instance iteration counts close to the single solver's."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MESHES = [(5, 3, 2), (17, 9, 4), (1, 1, 3), (16, 4, 8), (33, 2, 6), (12, 12, 10), (7, 7, 2), (20, 13, 5),
          (9, 11, 13), (6, 5, 16), (18, 10, 2)]


def _batch(sipg, meshes):
    dx = [m[0] for m in meshes]
    dy = [m[1] for m in meshes]
    ps = [m[2] for m in meshes]
    return sipg.Batch(dx, dy, ps, [sipg.pen_sigma(p) for p in ps])


@pytest.mark.parametrize("shape", [(16, 4), (8, 4), (8, 2)])
def test_batch_apply_equals_single_apply(sipg, shape):
    B = _batch(sipg, MESHES)
    for p in sorted({m[2] for m in MESHES}):
        rc = B.set_shape(p, *shape)
        assert rc in (sipg.OK, sipg.ERR_UNSUPPORTED)
        if rc == sipg.ERR_UNSUPPORTED:
            assert shape == (16, 4) and p >= 13
    x, y = B.vector(), B.vector()
    xs = []
    for i, (dx, dy, p) in enumerate(MESHES):
        xi = np.random.default_rng(10 + i).uniform(-1, 1, dx * dy * (p + 1) ** 2)
        xs.append(xi)
        B.upload(i, xi, x)
    assert B.apply(x, y) == sipg.OK
    for i, (dx, dy, p) in enumerate(MESHES):
        op = sipg.Operator(dx, dy, p, sipg.pen_sigma(p), sipg.F64)
        assert np.array_equal(B.download(i, y), op.apply_host(xs[i])), (i, shape)


def test_batch_tuning_picks_a_candidate(sipg):
    B = _batch(sipg, MESHES)
    B.tune()
    for p in sorted({m[2] for m in MESHES}):
        bx, by, us = B.tuning(p)
        assert (bx, by) in ((16, 4), (8, 4), (8, 2)) and us > 0
        assert not (p >= 13 and bx == 16)


def test_batch_tuning_is_cached_for_the_same_kind_of_batch(sipg):
    # an explicit tune records its choice; a second batch of the same
    # degrees and sizes tuned on first use takes it without timing (the
    # reported times are the recorded ones, to the bit)
    meshes = [(d, d, p) for d, p in [(12, 3), (20, 5), (16, 8), (10, 3)]]
    B1 = _batch(sipg, meshes)
    B1.tune()
    B2 = _batch(sipg, meshes)
    x, y = B2.vector(), B2.vector()
    x.zero()
    assert B2.apply(x, y) == sipg.OK            # first use: tunes
    for p in (3, 5, 8):
        assert B2.tuning(p) == B1.tuning(p), p


def _fourier_batch(sipg, meshes):
    B = _batch(sipg, meshes)
    b = B.vector()
    for i in range(len(meshes)):
        B.rhs_fourier(i, sipg.forcing_draw(i, 1, 4, 4), b)
    return B, b


SQUARE = [(4, 2), (8, 3), (16, 4), (24, 8), (5, 10), (32, 2), (12, 6), (20, 7), (9, 5), (4, 9), (28, 3), (16, 8)]


def test_batch_cg_reaches_tolerance_like_single_solves(sipg):
    meshes = [(d, d, p) for d, p in SQUARE]
    B, b = _fourier_batch(sipg, meshes)
    x = B.vector()
    x.zero()
    st, infos = B.cg(x, b, 1e-10, sipg.TOL_REL, replace_factor=3e-3)
    assert st == sipg.OK, sipg.strerror(st)
    for i, (d, _, p) in enumerate(meshes):
        op = sipg.Operator(d, d, p, sipg.pen_sigma(p), sipg.F64)
        xi = op.to_device(B.download(i, x))
        bi = op.to_device(B.download(i, b))
        rel = op.true_relres(xi, bi)
        assert rel <= 1e-10, (i, rel)
        assert infos[i].status == sipg.OK and infos[i].true_res_norm <= 1e-10 * infos[i].b_norm * 1.0000001
        xs = op.vector()
        xs.zero()
        st1, it1, _ = op.cg(xs, bi, 1e-10, sipg.TOL_REL)
        assert st1 == sipg.OK
        assert it1 - 3 <= infos[i].iters <= 1.25 * it1 + 64, (i, infos[i].iters, it1)
        ref = op.download(xs)
        got = B.download(i, x)
        assert np.linalg.norm(got - ref) <= 1e-7 * np.linalg.norm(ref), i


@pytest.mark.parametrize("shape", [(16, 4), (8, 4), (8, 2)])
def test_batch_cg_every_shape(sipg, shape):
    # 8x2 blocks of degree p have 16(p+1) threads: partial waves for every p
    meshes = [(d, d, p) for d, p in SQUARE[:8]]
    B, b = _fourier_batch(sipg, meshes)
    for p in sorted({m[2] for m in meshes}):
        assert B.set_shape(p, *shape) == sipg.OK
    x = B.vector()
    x.zero()
    st, infos = B.cg(x, b, 1e-10, sipg.TOL_REL, maxit=5000, replace_factor=3e-3)
    assert st == sipg.OK, [(k.status, k.iters) for k in infos]
    for i, (d, _, p) in enumerate(meshes):
        op = sipg.Operator(d, d, p, sipg.pen_sigma(p), sipg.F64)
        bi = op.to_device(B.download(i, b))
        xs = op.vector()
        xs.zero()
        st1, it1, _ = op.cg(xs, bi, 1e-10, sipg.TOL_REL)
        assert it1 - 3 <= infos[i].iters <= 1.25 * it1 + 64, (i, shape, infos[i].iters, it1)


def test_batch_skips_finished_instances_in_partial_waves(sipg):
    # p = 8 on 8x2 blocks: 144 threads, the last wave has 16 lanes; small
    # instances finish first, so the workgroups' item ranges cross finished
    # instances and every wave must skip to the same next item
    dims = [4, 30, 5, 26, 6, 20, 7, 17, 4, 24, 9, 28, 5, 19, 8, 22]
    meshes = [(d, d, 8) for d in dims]
    B, b = _fourier_batch(sipg, meshes)
    assert B.set_shape(8, 8, 2) == sipg.OK
    x = B.vector()
    x.zero()
    st, infos = B.cg(x, b, 1e-10, sipg.TOL_REL, maxit=5000, replace_factor=3e-3)
    assert st == sipg.OK, [(k.status, k.iters) for k in infos]
    for i in (1, 3, 11):
        d = dims[i]
        op = sipg.Operator(d, d, 8, sipg.pen_sigma(8), sipg.F64)
        bi = op.to_device(B.download(i, b))
        xs = op.vector()
        xs.zero()
        _, it1, _ = op.cg(xs, bi, 1e-10, sipg.TOL_REL)
        assert it1 - 3 <= infos[i].iters <= 1.25 * it1 + 64, (i, infos[i].iters, it1)


def test_batch_cg_deterministic(sipg):
    meshes = [(d, d, p) for d, p in SQUARE[:6]]
    B, b = _fourier_batch(sipg, meshes)
    x1, x2 = B.vector(), B.vector()
    x1.zero()
    x2.zero()
    st1, i1 = B.cg(x1, b, 1e-10, sipg.TOL_REL)
    st2, i2 = B.cg(x2, b, 1e-10, sipg.TOL_REL)
    assert st1 == st2 == sipg.OK
    assert [k.iters for k in i1] == [k.iters for k in i2]
    assert np.array_equal(x1.download(np.float64), x2.download(np.float64))


@pytest.mark.parametrize("method", ["cg", "pcg"])
def test_batch_without_graphs_is_the_same_solve(sipg, method):
    # direct launches of the captured kernels, in the same order: the same
    # iterates to the bit
    meshes = [(d, d, p) for d, p in SQUARE[:6]]
    B, b = _fourier_batch(sipg, meshes)
    x1, x2 = B.vector(), B.vector()
    x1.zero()
    x2.zero()
    st1, i1 = getattr(B, method)(x1, b, 1e-10, sipg.TOL_REL)
    B.set_graphs(False)
    st2, i2 = getattr(B, method)(x2, b, 1e-10, sipg.TOL_REL)
    assert st1 == st2 == sipg.OK
    assert [k.iters for k in i1] == [k.iters for k in i2]
    assert np.array_equal(x1.download(np.float64), x2.download(np.float64))


def test_batch_iteration_cap(sipg):
    meshes = [(d, d, p) for d, p in SQUARE[:5]]
    B, b = _fourier_batch(sipg, meshes)
    x = B.vector()
    x.zero()
    st, infos = B.cg(x, b, 1e-10, sipg.TOL_REL, maxit=5)
    assert st == sipg.NOT_CONVERGED
    assert all(k.iters == 5 and k.status == sipg.NOT_CONVERGED for k in infos)


def test_batch_breakdown_stays_in_its_instance(sipg):
    # a negative penalty makes one instance indefinite; the others converge
    meshes = [(8, 8, 3), (8, 8, 3), (12, 12, 4)]
    B = sipg.Batch([8, 8, 12], [8, 8, 12], [3, 3, 4], [80.0, -500.0, 125.0])
    b = B.vector()
    for i in range(3):
        B.rhs_fourier(i, sipg.forcing_draw(i, 1, 4, 4), b)
    x = B.vector()
    x.zero()
    st, infos = B.cg(x, b, 1e-10, sipg.TOL_REL, replace_factor=3e-3)
    assert st == sipg.ERR_BREAKDOWN
    assert infos[1].status == sipg.ERR_BREAKDOWN
    assert infos[0].status == sipg.OK and infos[2].status == sipg.OK
    del meshes


@pytest.mark.parametrize("shape", [None, (16, 4), (8, 4), (8, 2)])
def test_batch_pcg_matches_single_pcg(sipg, shape):
    # batched block-Jacobi PCG (sipg_batch_pcg) against sipg_pcg of each
    # instance alone: same tolerance on the true residual, a similar
    # iteration count, the same solution
    meshes = [(d, d, p) for d, p in SQUARE]
    B, b = _fourier_batch(sipg, meshes)
    if shape is not None:
        for p in sorted({m[2] for m in meshes}):
            assert B.set_shape(p, *shape) == sipg.OK
    x = B.vector()
    x.zero()
    st, infos = B.pcg(x, b, 1e-10, sipg.TOL_REL, maxit=5000, replace_factor=3e-3)
    assert st == sipg.OK, [(k.status, k.iters) for k in infos]
    for i, (d, _, p) in enumerate(meshes):
        op = sipg.Operator(d, d, p, sipg.pen_sigma(p), sipg.F64)
        xi = op.to_device(B.download(i, x))
        bi = op.to_device(B.download(i, b))
        assert op.true_relres(xi, bi) <= 1e-10, i
        xs = op.vector()
        xs.zero()
        # the recursion's own stop: the device ends it at the first
        # iteration below the tolerance, so the count is exact
        st1, info1 = op.pcg(xs, bi, 1e-10, sipg.TOL_REL, true_stop=False)
        assert st1 == sipg.OK
        assert info1.iters - 3 <= infos[i].iters <= 1.25 * info1.iters + 64, (i, infos[i].iters, info1.iters)
        ref, got = op.download(xs), B.download(i, x)
        assert np.linalg.norm(got - ref) <= 1e-7 * np.linalg.norm(ref), i


def test_batch_pcg_needs_the_standard_penalty(sipg):
    B = sipg.Batch([6, 6], [6, 6], [3, 3], [sipg.pen_sigma(3), 81.0])
    b, x = B.vector(), B.vector()
    b.zero()
    x.zero()
    st, _ = B.pcg(x, b, 1e-10, sipg.TOL_REL)
    assert st == 3          # SIPG_ERR_UNSUPPORTED


def test_batch_pcg_residual_swings_are_not_a_floor(sipg):
    # forcings whose PCG residual norm swings up to ~90x above its start
    # early on: the batched floor rules must leave them running
    meshes = [(256, 256, 8), (256, 256, 8)]
    B = _batch(sipg, meshes)
    b = B.vector()
    for i, seed in enumerate((1, 5)):
        B.rhs_fourier(i, sipg.forcing_draw(seed, 4, 4, 4), b)
    x = B.vector()
    x.zero()
    st, infos = B.pcg(x, b, 1e-8, sipg.TOL_REL, maxit=20000, replace_factor=1e-1)
    assert st == sipg.OK, [(k.status, k.iters) for k in infos]
