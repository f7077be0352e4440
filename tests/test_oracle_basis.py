"""Pins the oracle's basis numerics (LGL nodes/weights, derivative table D,
1-D stiffness K1) against closed forms, exactness properties, an
independent numpy route and the reference's own host code."""
import json
import math
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _ev(s):
    return eval(s, {"sqrt": math.sqrt})


@pytest.fixture(scope="module")
def thesis():
    with open(os.path.join(GOLD, "thesis_tables.json")) as f:
        return json.load(f)


@pytest.mark.parametrize("p", [2, 3, 4, 5, 6])
def test_gll_closed_forms(oracle, thesis, p):
    z, _ = oracle.gll(p)
    want = np.array([_ev(s) for s in thesis["gll_closed_forms"][str(p)]])
    assert np.max(np.abs(z - want)) < 5e-16


@pytest.mark.parametrize("p", list(range(1, 17)) + [24, 31])
def test_gll_vs_numpy_and_exactness(oracle, p):
    import dense
    z, w = oracle.gll(p)
    zn, wn = dense.gll_numpy(p)
    assert np.max(np.abs(z - zn)) < 1e-14
    assert np.max(np.abs(w - wn)) < 1e-14
    assert np.all(np.diff(z) > 0)
    assert np.max(np.abs(z + z[::-1])) < 1e-15          # symmetric
    # exact for polynomials up to degree 2p-1 (Thesis p.40 Table 4.1)
    for d in range(0, 2 * p):
        exact = 0.0 if d % 2 else 2.0 / (d + 1)
        assert abs(np.sum(w * z ** d) - exact) < 1e-13
    # and not exact at degree 2p (a Lobatto rule of p+1 points)
    if p <= 12:
        assert abs(np.sum(w * z ** (2 * p)) - 2.0 / (2 * p + 1)) > 1e-11


@pytest.mark.parametrize("p", range(1, 17))
def test_dphi_differentiates_polynomials(oracle, p):
    z, _ = oracle.gll(p)
    D = oracle.dphi(p)     # D[mode, node]
    for k in range(0, p + 1):
        # sum_mode z_mode^k phi'_mode(z_node) = k z_node^(k-1)
        got = (z ** k) @ D
        want = k * z ** (k - 1) if k > 0 else np.zeros_like(z)
        assert np.max(np.abs(got - want)) < 1e-12 * max(1, p ** 2)


@pytest.mark.parametrize("p", [2, 3])
def test_k1_closed_forms(oracle, thesis, p):
    K = oracle.k1(p)
    want = np.array([[_ev(s) for s in row] for row in thesis["k1_closed_forms"][str(p)]])
    assert np.max(np.abs(K - want)) < 1e-14


@pytest.mark.parametrize("p", range(1, 17))
def test_k1_identity(oracle, p):
    z, w = oracle.gll(p)
    D = oracle.dphi(p)
    K = oracle.k1(p)
    assert np.allclose(K, (D * w) @ D.T, rtol=0, atol=1e-12 * p * p)
    assert np.max(np.abs(K - K.T)) < 1e-13 * p * p
    assert np.max(np.abs(K @ np.ones(p + 1))) < 1e-12 * p * p      # kills constants


def test_basis_matches_reference_golden(oracle):
    with open(os.path.join(GOLD, "basis_ref.json")) as f:
        ref = json.load(f)["p"]
    assert {20, 24, 32, 40, 48, 63} <= set(map(int, ref))
    for p in sorted(map(int, ref)):
        r = ref[str(p)]
        z, w = oracle.gll(p)
        D = oracle.dphi(p)
        K = oracle.k1(p)
        for name, mine in (("nodes", z), ("weights", w), ("D", D), ("K1", K)):
            theirs = np.array([float.fromhex(s) for s in r[name]])
            scale = max(1.0, np.max(np.abs(theirs)))
            assert np.max(np.abs(mine.ravel() - theirs)) <= 4e-15 * scale * (p + 1), (p, name)


@pytest.mark.skipif(not os.path.exists(os.path.join(os.path.dirname(__file__), "..", "oracle", "_ref",
                                                    "libref_basis.so")),
                    reason="reference host numerics not built (make -C oracle ref)")
@pytest.mark.parametrize("p", [1, 5, 9, 20, 31])
def test_basis_matches_reference_host_code(oracle, p):
    z, w = oracle.gll(p)
    zr, wr, Dr, Kr = oracle.ref_basis(p)
    assert np.max(np.abs(z - zr)) < 1e-15 * (p + 1)
    assert np.max(np.abs(w - wr)) < 1e-15 * (p + 1)
    scale = np.max(np.abs(Dr))
    assert np.max(np.abs(oracle.dphi(p) - Dr)) < 1e-14 * scale * (p + 1)
