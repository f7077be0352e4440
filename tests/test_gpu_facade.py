"""The C++ facade (include/sipg_hip.hpp): a compiled C++ program written
against the reference's API (sipg_sem_2d, abs_mvm, conjugate_gradient,
preconditioned_conjugate_gradient, mode_vector, host_mode_vector) and
linked to libsipg_hip.so alone solves configuration C1 (8x8, p = 3, pen
80, seeded Fourier forcing) and reproduces the dense LAPACK solution, with
CG and with PCG, through the library's solvers and through the generic
loops over a user abs_mvm; and reproduces the thesis' Table 4.10 rows
(solution E) through sipg_sem_2d."""
import json
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CPP = os.path.join(HERE, "cpp")
GOLD = os.path.join(HERE, "golden")


def _build():
    subprocess.run(["make", "-s", "-C", CPP], check=True, timeout=300)


def test_facade_compiles_against_the_header_and_library():
    # host compile and link only (no GPU needed)
    _build()
    for tag in ("cg", "pcg"):
        assert os.access(os.path.join(CPP, "_build", f"facade_test_{tag}"), os.X_OK)


@pytest.fixture(scope="module", params=["cg", "pcg"])
def run(request, tmp_path_factory):
    _build()
    g = np.load(os.path.join(GOLD, "c1_solution.npz"))
    out = tmp_path_factory.mktemp(f"facade_{request.param}")
    args = [os.path.join(CPP, "_build", f"facade_test_{request.param}"), str(out)]
    for a, m, n in zip(g["a"], g["m"], g["n"]):
        args += [repr(float(a)), str(int(m)), str(int(n))]
    print(f"[facade] {' '.join(args)}", flush=True)
    p = subprocess.run(args, capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr
    cases = {r["case"]: r for r in map(json.loads, p.stdout.strip().splitlines())}
    return request.param, out, cases, g


def _load(out, name):
    return np.fromfile(os.path.join(out, name), dtype=np.float64)


@pytest.mark.gpu
def test_sem_2d_reproduces_the_lapack_solution(run):
    tag, out, cases, g = run
    u = _load(out, f"c1_{tag}_u.bin")
    rhs = _load(out, f"c1_{tag}_rhs.bin")
    assert np.max(np.abs(rhs - g["rhs"])) <= 1e-13 * np.abs(g["rhs"]).max()
    assert np.linalg.norm(u - g["u"]) <= 1e-9 * np.linalg.norm(g["u"])
    r = cases[f"c1_sem_{tag}"]
    assert r["iterations"] > 0
    # the discretisation error of C1 against the Fourier solution
    assert 0 < r["L2_err"] < r["H1_err"] and r["max_err"] < 1e-2


@pytest.mark.gpu
@pytest.mark.parametrize("solver", ["native_cg", "native_pcg", "generic_cg", "generic_pcg"])
def test_free_solvers_reproduce_the_lapack_solution(run, solver):
    tag, out, cases, g = run
    u = _load(out, f"c1_{tag}_{solver}.bin")
    assert np.linalg.norm(u - g["u"]) <= 1e-9 * np.linalg.norm(g["u"]), solver
    its = cases[f"c1_{tag}_{solver}"]["iterations"]
    assert 0 < its < 1024
    if solver.endswith("pcg"):
        assert its < cases[f"c1_{tag}_{solver[:-3]}cg"]["iterations"]


@pytest.mark.gpu
def test_mode_vector_round_trip(run):
    tag, _, cases, _ = run
    r = cases[f"round_trip_{tag}"]
    assert r["same"] and r["element_major"], r


@pytest.mark.gpu
def test_sem_2d_in_fp32(run):
    tag, out, cases, g = run
    u = _load(out, f"c1_{tag}_u_f32.bin")
    assert np.linalg.norm(u - g["u"]) <= 1e-4 * np.linalg.norm(g["u"])


@pytest.mark.gpu
def test_sem_2d_reproduces_thesis_table_4_10(run):
    tag, _, cases, _ = run
    with open(os.path.join(GOLD, "thesis_tables.json")) as f:
        rows = json.load(f)["solutions"]["E"]["rows"]
    n = 0
    for r in rows:
        c = cases.get(f"E_{tag}_{r['dim']}_{r['p']}")
        if c is None:
            continue
        n += 1
        for key, ours in (("H1", c["H1_err"]), ("L2", c["L2_err"]), ("Linf", c["max_err"])):
            assert abs(ours - r[key]) <= 2e-5 * r[key], (r, key, ours)
    assert n == 8
