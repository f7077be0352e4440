"""Pins the oracle's matrix-free operator (the reference's volume +
flux_term6a + flux_term6b restated) against an independent dense assembly
of the SIPG bilinear form, and checks its structural invariants."""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def matrix_of(oracle, dx, dy, p, pen):
    n = dx * dy * (p + 1) ** 2
    M = np.zeros((n, n))
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0
        M[:, j] = oracle.apply(dx, dy, p, pen, e)
    return M


@pytest.mark.parametrize("dx,dy,p,pen", [(1, 1, 1, 3.0), (1, 1, 2, 7.0), (2, 2, 2, 45.0),
                                         (3, 2, 3, 80.0), (2, 3, 4, 125.0), (1, 4, 2, 9.0),
                                         (4, 1, 3, 16.0), (3, 3, 1, 20.0)])
def test_apply_equals_dense_bilinear_form(oracle, dx, dy, p, pen):
    import dense
    A = dense.assemble(dx, dy, p, pen)
    M = matrix_of(oracle, dx, dy, p, pen)
    assert np.max(np.abs(M - A)) <= 1e-13 * np.max(np.abs(A))


def test_dense_fixture(oracle):
    A = np.load(os.path.join(GOLD, "dense_d2p2.npy"))
    M = matrix_of(oracle, 2, 2, 2, 45.0)
    assert np.max(np.abs(M - A)) <= 1e-13 * np.max(np.abs(A))


@pytest.mark.parametrize("dx,dy,p,pen", [(3, 2, 3, 80.0), (2, 2, 5, 180.0)])
def test_spd_above_coercivity(oracle, dx, dy, p, pen):
    M = matrix_of(oracle, dx, dy, p, pen)
    assert np.max(np.abs(M - M.T)) <= 1e-13 * np.max(np.abs(M))
    assert np.linalg.eigvalsh(0.5 * (M + M.T)).min() > 0


def test_indefinite_without_penalty(oracle):
    # pen = 0 is below coercivity (Thesis p.28 Lemma 2.12): A is indefinite
    M = matrix_of(oracle, 2, 2, 3, 0.0)
    assert np.linalg.eigvalsh(0.5 * (M + M.T)).min() < 0


def test_volume_kills_constants_inside(oracle):
    # on an interior element the volume and face terms of a constant cancel
    dx = dy = 3
    p = 4
    n = dx * dy * (p + 1) ** 2
    y = oracle.apply(dx, dy, p, 50.0, np.ones(n))
    noe = dx * dy
    centre = 4
    assert np.max(np.abs(y.reshape(-1, noe)[:, centre])) < 1e-12


@pytest.mark.parametrize("name", ["d4p3", "d3p8", "d2p15", "r5x3p2"])
def test_apply_vectors_fixture(oracle, name):
    g = np.load(os.path.join(GOLD, "apply_vectors.npz"))
    dx, dy, p, pen = g[name + "_shape"]
    y = oracle.apply(int(dx), int(dy), int(p), float(pen), g[name + "_x"])
    ref = g[name + "_y"]
    assert np.max(np.abs(y - ref)) <= 1e-14 * np.max(np.abs(ref))


def test_apply_f32_close_to_f64(oracle):
    rng = np.random.default_rng(3)
    x = rng.uniform(-1, 1, 4 * 4 * 36)
    y64 = oracle.apply(4, 4, 5, 180.0, x)
    y32 = oracle.apply_f32(4, 4, 5, 180.0, x.astype(np.float32))
    assert np.max(np.abs(y32 - y64)) <= 2e-6 * np.max(np.abs(y64))
