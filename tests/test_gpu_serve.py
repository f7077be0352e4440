"""The serving loop (sipg_server_*): requests admitted within an HBM budget,
their CG state resident on the device, several in flight at once.  Each
result must be bit for bit the one the same request gets alone, and meet
its tolerance on the true residual of the single-problem operator."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _requests(n):
    rng = np.random.default_rng(7)
    out = []
    for k in range(n):
        dim = int(rng.integers(4, 25))
        p = int(rng.integers(2, 9))
        tol = 1e-10 if rng.random() < 0.5 else 1e-8
        out.append((dim, p, tol, k))
    return out


def _serve(sipg, reqs, budget):
    S = sipg.Server(budget)
    ids = {}
    for i, (dim, p, tol, seed) in enumerate(reqs):
        ids[S.submit(dim, p, sipg.pen_sigma(p), sipg.forcing_draw(seed, 1, 4, 4), tol)] = i
    out, max_queued = {}, 0
    for _ in range(20000):
        S.progress()
        max_queued = max(max_queued, S.stats()["queued"])
        while True:
            r = S.take(64 * 64 * 121)
            if r is None:
                break
            i = ids[r[0]]
            dim, p = reqs[i][0], reqs[i][1]
            out[i] = (r[1], r[2][: dim * dim * (p + 1) ** 2].copy())
        if len(out) == len(reqs):
            break
    return out, max_queued


def test_server_results_alone_and_in_the_mix(sipg):
    reqs = _requests(16)
    demand = sum(48 * d * d * (p + 1) ** 2 for d, p, _, _ in reqs)
    budget = max(demand / 3, max(48 * d * d * (p + 1) ** 2 for d, p, _, _ in reqs))
    mix, max_queued = _serve(sipg, reqs, budget)
    assert len(mix) == len(reqs) and max_queued > 0
    for i, (dim, p, tol, seed) in enumerate(reqs):
        info, x = mix[i]
        assert info.status == sipg.OK, (i, info.status)
        alone, _ = _serve(sipg, [reqs[i]], budget)
        assert np.array_equal(alone[0][1], x), i
        op = sipg.Operator(dim, dim, p, sipg.pen_sigma(p), sipg.F64)
        b = op.vector()
        op.rhs_fourier(sipg.forcing_draw(seed, 1, 4, 4), b)
        rel = op.true_relres(op.to_device(x), b)
        assert rel <= tol * 1.0000001, (i, rel, tol)


def test_server_rejects_a_request_over_budget(sipg):
    S = sipg.Server(1e6)
    with pytest.raises(sipg.SipgError) as e:
        S.submit(64, 10, sipg.pen_sigma(10), sipg.forcing_draw(0, 1, 4, 4), 1e-8)
    assert e.value.status == sipg.ERR_BUDGET
