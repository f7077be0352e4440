"""The batched (C4) and served (C5) configurations at their workload shape,
checked against the oracle (the reference's CG restated in C, run to a
1e-13 relative residual).

C4: 2048 independent FP64 solves in one batch, built exactly as the
benchmark builds it (mesh side ~ U{4..64}, degree ~ U{2..10} from numpy
default_rng(4), instance i forced by forcing_draw(i, 1..4 modes, m,n <= 4),
pen = 5 (p+1)^2), each to a 1e-10 true relative residual by batched PCG
and by batched CG; every instance's true residual is checked, and every
128th instance's solution against the oracle.

C5: a request mix of the serving configuration's shape (mesh side
log-uniform, degree ~ U{2..12}, 70 % marked mixed precision, tolerance
1e-8 or 1e-10 with equal probability), mesh sides limited to 8..64 so the
oracle can check every request.  Each served solution is compared with the
oracle's solve of the same request; the measured differences are printed.
A second mix of large meshes near their FP64 floor (sides 150-250, degree
8-12, tolerance 1e-10) must stop well below the iteration cap."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
TOL = 1e-10
N4 = 2048


def _c4(sipg):
    rng = np.random.default_rng(4)
    dims = rng.integers(4, 65, N4)
    ps = rng.integers(2, 11, N4)
    pens = [sipg.pen_sigma(int(p)) for p in ps]
    B = sipg.Batch(dims, dims, ps, pens)
    b = B.vector()
    for i in range(N4):
        B.rhs_fourier(i, sipg.forcing_draw(i, 1, 4, 4), b)
    B.tune()
    return B, b, dims, ps, pens


@pytest.fixture(scope="module")
def c4(sipg):
    B, b, dims, ps, pens = _c4(sipg)
    out = {"B": B, "b": b, "dims": dims, "ps": ps, "pens": pens}
    for method in ("pcg", "cg"):
        x = B.vector()
        x.zero()
        solve = B.pcg if method == "pcg" else B.cg
        st, infos = solve(x, b, TOL, sipg.TOL_REL, maxit=20000, replace_factor=3e-3)
        out[method] = (st, infos, x)
    return out


@pytest.mark.parametrize("method", ["pcg", "cg"])
def test_c4_every_instance_meets_the_tolerance(sipg, c4, method):
    st, infos, _ = c4[method]
    assert st == sipg.OK, sipg.strerror(st)
    bad = [(i, k.status, k.true_res_norm / k.b_norm) for i, k in enumerate(infos)
           if k.status != sipg.OK or not 0.0 <= k.true_res_norm <= TOL * k.b_norm * (1 + 1e-12)]
    assert not bad, bad[:10]


@pytest.mark.parametrize("i", list(range(0, N4, 128)))
def test_c4_sample_matches_the_oracle(sipg, oracle, c4, i):
    B, b = c4["B"], c4["b"]
    d, p, pen = int(c4["dims"][i]), int(c4["ps"][i]), c4["pens"][i]
    bh = B.download(i, b)
    xo, _, _ = oracle.cg(d, d, p, pen, bh, tol=1e-13, tol_mode=1)
    for method in ("pcg", "cg"):
        xh = B.download(i, c4[method][2])
        rel = np.linalg.norm(xh - xo) / np.linalg.norm(xo)
        print(f"C4 instance {i} ({d}x{d}, p={p}) {method}: rel L2 vs oracle {rel:.2e}")
        assert rel <= 1e-8, (method, rel)


def _l2_fn(dim, p, u):
    """L2 norm of a nodal function (reference order) by GLL quadrature, as
    sipg_sem_2d::err_norms (sipg_sem_2d.hpp:336-360)."""
    from sipg import basis_tables
    _, w, _, _ = basis_tables(p)
    h = 1.0 / dim
    W = np.outer(w, w).reshape(-1) * (h / 2) ** 2
    u = u.reshape((p + 1) ** 2, dim * dim)
    return float(np.sqrt(np.sum(W[:, None] * u * u)))


def _c5_requests(n, dmin, dmax, seed=0x5EED):
    rng = np.random.default_rng(seed)
    dims = np.rint(np.exp(rng.uniform(np.log(dmin), np.log(dmax), n))).astype(int)
    ps = rng.integers(2, 13, n)
    mixed = rng.random(n) < 0.7
    tols = np.where(rng.random(n) < 0.5, 1e-8, 1e-10)
    return [(int(dims[k]), int(ps[k]), float(tols[k]), bool(mixed[k])) for k in range(n)]


def _serve(sipg, reqs, budget, maxit=30000):
    S = sipg.Server(budget)
    assert S.set_method(sipg.METHOD_PCG) == sipg.OK
    ids = {}
    for k, (dim, p, tol, _) in enumerate(reqs):
        ids[S.submit(dim, p, sipg.pen_sigma(p), sipg.forcing_draw(k, 1, 4, 4), tol, maxit=maxit)] = k
    out = {}
    nmax = max(d * d * (p + 1) ** 2 for d, p, _, _ in reqs)
    for _ in range(100000):
        S.progress()
        while True:
            r = S.take_ex(nmax)
            if r is None:
                break
            k = ids[r[0]]
            d, p = reqs[k][0], reqs[k][1]
            out[k] = (r[1], r[2][: d * d * (p + 1) ** 2].copy(), r[3])
        if len(out) == len(reqs):
            break
    return out


def test_c5_shaped_mix_matches_the_oracle(sipg, oracle):
    reqs = _c5_requests(24, 8, 64)
    demand = sum(56 * d * d * (p + 1) ** 2 for d, p, _, _ in reqs)
    out = _serve(sipg, reqs, max(demand / 3, 56 * 64 * 64 * 169))
    assert len(out) == len(reqs)
    worst = {1e-8: 0.0, 1e-10: 0.0}
    for k, (d, p, tol, mixed) in enumerate(reqs):
        info, xh, t = out[k]
        assert info.status == sipg.OK, (k, info.status)
        assert 0.0 <= info.true_res_norm <= tol * info.b_norm * (1 + 1e-12), k
        assert 0.0 <= t["submitted"] <= t["admitted"] <= t["completed"], t
        pen = sipg.pen_sigma(p)
        op = sipg.Operator(d, d, p, pen, sipg.F64)
        b = op.vector()
        op.rhs_fourier(sipg.forcing_draw(k, 1, 4, 4), b)
        xo, _, _ = oracle.cg(d, d, p, pen, op.download(b), tol=1e-13, tol_mode=1)
        l2 = _l2_fn(d, p, xh - xo) / _l2_fn(d, p, xo)
        worst[tol] = max(worst[tol], l2)
        print(f"C5 request {k} ({d}x{d}, p={p}, tol {tol:.0e}, marked {'mixed' if mixed else 'fp64'}): "
              f"{info.iters} iterations, relative L2 difference vs the oracle {l2:.2e}")
    print(f"worst relative L2 difference: tol 1e-10 {worst[1e-10]:.2e}, tol 1e-8 {worst[1e-8]:.2e}")
    assert worst[1e-10] <= 1e-9
    assert worst[1e-8] <= 1e-6


def test_c5_large_meshes_stop_well_below_the_cap(sipg):
    # sides 150-250, degree 8-12, tolerance 1e-10: near the FP64 floor of
    # these meshes; each request must end (converged, or stopped at the
    # floor) long before the 30,000-iteration cap
    reqs = [(150, 8, 1e-10, False), (200, 10, 1e-10, False), (230, 12, 1e-10, False), (250, 9, 1e-10, False)]
    out = _serve(sipg, reqs, 80e9)
    for k, (d, p, tol, _) in enumerate(reqs):
        info, _, t = out[k]
        rel = info.true_res_norm / info.b_norm
        print(f"large request {d}x{d} p={p}: {sipg.strerror(info.status)}, {info.iters} iterations, "
              f"true relres {rel:.2e}, service {t['completed'] - t['admitted']:.2f} s")
        assert info.status in (sipg.OK, sipg.NOT_CONVERGED)
        assert info.iters < 10000, info.iters
        assert rel <= 1e-9


def test_server_cancel_and_timeout_on_the_device(sipg):
    reqs = _c5_requests(6, 8, 32, seed=3)
    S = sipg.Server(56 * 40 * 40 * 169)
    assert S.set_method(sipg.METHOD_PCG) == sipg.OK
    ids = [S.submit(d, p, sipg.pen_sigma(p), sipg.forcing_draw(k, 1, 4, 4), 1e-12, maxit=100000)
           for k, (d, p, _, _) in enumerate(reqs)]
    assert S.cancel(ids[-1]) == sipg.OK                  # queued
    S.progress()
    assert S.cancel(ids[0]) == sipg.OK                   # running
    assert S.cancel(123456789) == sipg.ERR_NOT_FOUND
    S.set_timeout(1e-9)                                  # everything else is overdue
    got = {}
    for _ in range(100):
        S.progress()
        while (r := S.take_ex(64 * 64 * 169)) is not None:
            got[r[0]] = r
        if len(got) == len(ids):
            break
    assert len(got) == len(ids)
    assert all(got[i][1].status == sipg.ERR_CANCELLED for i in ids)
    assert got[ids[-1]][3]["admitted"] == -1.0 and got[ids[-1]][1].iters == 0
    assert got[ids[0]][1].iters > 0
