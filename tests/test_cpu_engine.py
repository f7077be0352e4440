"""The solver engine's host orchestration (abi.cpp, solvers.cpp, basis.cpp
-- the shipped sources) compiled for the CPU against a mock HIP runtime
and CPU stand-ins of the kernel launchers, under AddressSanitizer and
UBSan.  Exercises the device-state protocol (parity slots, carry-forward,
residual replacement, iteration cap, mixed-precision folding) without a
GPU; the numerics are the oracle's."""
import json
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ENG = os.path.join(HERE, "cpu_engine")


@pytest.fixture(scope="module")
def results():
    if shutil.which("g++") is None:
        pytest.skip("no host compiler")
    subprocess.run(["make", "-s", "-C", ENG], check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0", UBSAN_OPTIONS="halt_on_error=1")
    out = subprocess.run([os.path.join(ENG, "_build", "engine_driver")], check=True,
                         capture_output=True, text=True, env=env, timeout=600)
    assert "runtime error" not in out.stderr and "AddressSanitizer" not in out.stderr, out.stderr
    return {r["case"]: r for r in map(json.loads, out.stdout.strip().splitlines())}


# The engine's vectors are element-major and the oracle's are in the
# reference order, so dot products sum the same terms in another order:
# the iterates agree to rounding, not bit for bit.
ROUNDING = 1e-11


def test_plain_cg_is_the_reference_iteration(results):
    r = results["plain"]
    assert r["rc"] == 0 and r["iters"] == r["oracle_iters"] and r["diff"] <= ROUNDING


def test_iteration_cap(results):
    r = results["plain_cap"]
    assert r["rc"] == 6 and r["iters"] == 5 and r["diff"] <= ROUNDING


# With a true-residual stop the device keeps iterating after the recursive
# residual meets the tolerance until the host has inserted the true
# residual; the host polls one chunk (at most 8 iterations near the
# tolerance, 32 elsewhere) behind.
LAG = 48


def test_replacement_reaches_true_residual(results):
    # reliable updates keep the search direction: the iteration count stays
    # that of the reference iteration, plus the polling lag
    r = results["ex_replace"]
    assert r["rc"] == 0 and r["reps"] >= 2 and r["relres"] <= 1e-10
    assert r["oracle_iters"] <= r["iters"] <= r["oracle_iters"] + LAG


def test_true_residual_stop_without_periodic_replacement(results):
    r = results["ex_noreplace"]
    assert r["rc"] == 0 and r["reps"] == 1 and r["relres"] <= 1e-10
    assert r["oracle_iters"] <= r["iters"] <= r["oracle_iters"] + LAG


@pytest.mark.parametrize("case", ["mixed", "mixed_keep_p", "pcg_mixed"])
def test_mixed_precision_converges_in_fp64(results, case):
    r = results[case]
    assert r["rc"] == 0 and r["relres"] <= 1e-10 and r["diff"] <= 1e-8


def test_mixed_cap_reports_not_converged(results):
    r = results["mixed_cap"]
    assert r["rc"] == 6 and r["iters"] == 30


@pytest.mark.parametrize("case", ["pcg_replace", "pcg_true_stop", "pcg_other_pen"])
def test_pcg_reaches_true_residual_in_fewer_iterations(results, case):
    # block-Jacobi PCG (the host protocol of sipg_pcg with the kernels'
    # CPU stand-ins): same solution as the oracle's CG, fewer iterations
    r = results[case]
    assert r["rc"] == 0 and r["relres"] <= 1e-10 and r["diff"] <= 1e-8, r
    assert r["iters"] < 0.8 * r["oracle_iters"], r


@pytest.mark.parametrize("case", ["coarse_tables_5x5_p3", "coarse_tables_6x4_p2", "coarse_tables_4x5_p1"])
def test_coarse_tables_reproduce_the_galerkin_coarse_operator(results, case):
    # the sine-transform tables of the two-level preconditioner against
    # P0 A P0^T built from the oracle's operator and bilinear interpolation
    # (square, rectangular, and p = 1 with its lumped GLL mass)
    r = results[case]
    assert r["rel"] < 1e-13, r
    assert abs(r["center"] - (4.0 if case.endswith("p1") else 8.0 / 3.0)) < 1e-6, r      # printed to 6 places


@pytest.mark.parametrize("case,n", [("pcg_floor", 1600), ("pcg2_floor", 2304)])
def test_preconditioned_solvers_stop_at_the_fp64_floor(results, case, n):
    # tolerance 1e-17, below what FP64 attains: the recursion stalls at the
    # last true residual and the solve ends there, not converged, well
    # before its iteration cap (n) -- with the true residual at the floor
    r = results[case]
    assert r["rc"] == 6 and r["iters"] < 0.7 * n, r
    assert r["relres"] < 1e-13, r


@pytest.mark.parametrize("case", ["ex_near_floor", "ex_nearer_floor", "pcg_near_floor", "pcg2_near_floor"])
def test_near_the_floor_a_peek_can_finish_the_solve(results, case):
    # tolerances just above this problem's FP64 floor (~1.26e-14): 2e-14,
    # and 1.3e-14 for CG.  The finishing checks need the recursion to fall
    # 4x below the last true residual, which takes long there; the true
    # residuals computed in between without replacing r (peeks) end the
    # solve on the first that meets the tolerance.  The solution's true
    # residual, recomputed independently, meets it.
    tol = 1.3e-14 if case == "ex_nearer_floor" else 2e-14
    r = results[case]
    assert r["rc"] == 0 and r["relres"] <= tol, r
    assert r["true_res"] <= tol * r["b_norm"], r


@pytest.mark.parametrize("case", ["ex_replace", "pcg_replace", "pcg_true_stop", "pcg2_replace", "pcg2_true_stop",
                                  "pcg_near_floor", "pcg2_near_floor"])
def test_true_residual_stops_queue_little_past_the_finish(results, case):
    # the engine polls the solve state a chunk behind the device; iterations
    # queued past the finish are no-ops that still cost their launches.  The
    # finish is predicted from the fall of r.r, chunks end short of it, and
    # near it the host waits for the newest state, so that with a
    # true-residual stop at most two iterations are enqueued past the
    # reported count (none in these cases)
    r = results[case]
    assert r["rc"] == 0, r
    assert r["launched"] - r["iters"] <= 2, r


def test_near_the_floor_cg_finishes_by_a_peek(results):
    # which of the two CG tolerances a peek ends depends on where the host's
    # polls fall; at least one does
    peeks = [results[c]["peeks"] for c in ("ex_near_floor", "ex_nearer_floor")]
    assert max(peeks) >= 1, peeks


def test_plain_cg_at_the_floor_runs_to_its_cap(results):
    # plain CG's residual swings too far for the stall test: it reports the
    # cap, with the true residual at the floor
    r = results["ex_floor"]
    assert r["rc"] == 6 and r["iters"] == 1600 and r["relres"] < 1e-13, r


@pytest.mark.parametrize("case", ["problem_1_4_p2", "problem_2_5_p3", "problem_3_3_p4", "problem_4_6_p2",
                                  "problem_5_8_p3", "problem_5_1_p5"])
def test_general_rhs_assembly_is_the_reference_loop(results, case):
    # sipg_rhs_nodal fed with f and g = u_ex at sipg_quad_points: the host
    # assembly (problem.cpp) reproduces orc_rhs, which follows the
    # reference's compute_rhs term for term, to the bit; the nodal norms'
    # host side (partials, reduction) against orc_err_norms
    r = results[case]
    assert r["rc"] == 0 and r["rhs_rel"] == 0.0, r
    assert r["norms_rel"] < 1e-13, r


@pytest.mark.parametrize("case", ["pcg2_replace", "pcg2_true_stop", "pcg2_plain", "pcg2_one_element"])
def test_pcg2_converges_to_the_oracle_solution(results, case):
    r = results[case]
    assert r["rc"] == 0 and r["relres"] <= 1e-10 and r["diff"] < 1e-9, r


def test_pcg2_needs_far_fewer_iterations(results):
    two, one = results["pcg2_plain"], results["pcg2_vs_pcg"]
    assert two["iters"] < 0.4 * one["iters"], (two, one)


def test_pcg_plain_stops_on_the_recursion(results):
    r = results["pcg_plain"]
    assert r["rc"] == 0 and r["reps"] == 0 and r["relres"] <= 2e-10, r
    assert r["iters"] < 0.8 * r["oracle_iters"], r


def test_pcg_iteration_cap(results):
    r = results["pcg_cap"]
    assert r["rc"] == 6 and r["iters"] == 5, r


@pytest.fixture(scope="module")
def batch_results(results):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0", UBSAN_OPTIONS="halt_on_error=1")
    out = subprocess.run([os.path.join(ENG, "_build", "batch_driver")], check=True, capture_output=True,
                         text=True, env=env, timeout=600)
    assert "runtime error" not in out.stderr and "AddressSanitizer" not in out.stderr, out.stderr
    rows = [json.loads(l) for l in out.stdout.strip().splitlines()]
    by = {}
    for r in rows:
        by.setdefault(r["case"], []).append(r)
    return by


def test_batch_engine_solves_every_instance(batch_results):
    # graph capture and replay, list compaction and true-residual checks of
    # the batched engine (batch.cpp); each instance is the reference CG
    for r in batch_results["batch"]:
        assert r["rc"] == 0 and r["status"] == 0 and r["relres"] <= 1e-10, r
        # the finish is confirmed on a true residual inserted a replay or
        # two (8 iterations each) after the recursion met the tolerance
        assert r["oracle_iters"] - 2 <= r["iters"] <= r["oracle_iters"] + 40, r
        assert r["diff"] <= 1e-9 and r["reps"] >= 1 and r["apply_err"] == 0.0, r


def test_batch_engine_iteration_cap(batch_results):
    for r in batch_results["batch_cap"]:
        assert r["rc"] == 6 and r["status"] == 6 and r["iters"] == 5, r


def test_batch_engine_unbounded_maxit(batch_results):
    # maxit = INT_MAX: every instance converges as with the default cap
    for r in batch_results["batch_maxit_intmax"]:
        assert r["rc"] == 0 and r["status"] == 0 and r["relres"] <= 1e-10, r


def test_batch_engine_breakdown_is_per_instance(batch_results):
    rows = batch_results["batch_breakdown"]
    assert all(r["rc"] == 8 for r in rows)
    assert [r["status"] for r in rows] == [0, 8, 0]
    assert rows[0]["relres"] <= 1e-10 and rows[2]["relres"] <= 1e-10


def test_batch_pcg_solves_every_instance_in_fewer_iterations(batch_results):
    # batched block-Jacobi PCG: same solutions as the oracle's CG, fewer
    # iterations for every instance but the smallest
    for r in batch_results["batch_pcg"]:
        assert r["rc"] == 0 and r["status"] == 0 and r["relres"] <= 1e-10, r
        assert r["diff"] <= 1e-8 and r["iters"] <= r["oracle_iters"] + 16, r
    its = sum(r["iters"] for r in batch_results["batch_pcg"])
    ito = sum(r["oracle_iters"] for r in batch_results["batch_pcg"])
    assert its < 0.8 * ito


def test_batch_pcg_needs_the_standard_penalty(batch_results):
    assert all(r["rc"] == 3 for r in batch_results["batch_pcg_other_pen"])


@pytest.fixture(scope="module")
def serve_results(results):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0", UBSAN_OPTIONS="halt_on_error=1")
    out = subprocess.run([os.path.join(ENG, "_build", "serve_driver")], check=True, capture_output=True,
                         text=True, env=env, timeout=600)
    assert "runtime error" not in out.stderr and "AddressSanitizer" not in out.stderr, out.stderr
    return [json.loads(l) for l in out.stdout.strip().splitlines()]


def test_batch_and_server_under_deferred_execution(batch_results, serve_results):
    # both drivers run with the mock's deferred model: copies, launches and
    # graph replays execute only at the engine's synchronisation points and
    # read host buffers and graphs then, under ASan/UBSan (the fixtures
    # fail on any report); every queued item has run once the calls return
    assert batch_results["deferred_queue"][0]["pending"] == 0
    q = [r for r in serve_results if r["case"] == "deferred_queue"]
    assert q and q[0]["pending"] == 0


def test_server_admits_within_budget_and_finishes_all(serve_results):
    mix = serve_results[0]
    # the budget holds about a third of the demand: requests had to wait
    assert mix["case"] == "serve_mix" and mix["rc"] == 0 and mix["done"] == mix["total"] == 24
    assert mix["max_queued"] > 0


def test_server_result_same_alone_and_in_the_mix(serve_results):
    ones = [r for r in serve_results if r["case"] == "serve_one"]
    assert len(ones) == 24
    for r in ones:
        assert r["rc"] == 0 and r["status"] == 0, r
        assert 0 <= r["true_rel"] <= r["tol"], r
        assert r["alone_maxdiff"] == 0.0, r


def test_server_pcg_finishes_all_like_cg(serve_results):
    mix = next(r for r in serve_results if r["case"] == "serve_pcg_mix")
    assert mix["rc"] == 0 and mix["done"] == mix["total"] == 24 and mix["max_queued"] > 0
    ones = [r for r in serve_results if r["case"] == "serve_pcg_one"]
    assert len(ones) == 24
    for r in ones:
        assert r["status"] == 0 and 0 <= r["true_rel"] <= r["tol"], r
        # two solutions to the same tolerance: their distance is of the
        # order of the tolerance times the condition number's growth
        assert r["rel_vs_cg"] <= 1e4 * r["tol"], r
        if r["alone_maxdiff"] >= 0:
            assert r["alone_maxdiff"] == 0.0, r
    assert sum(r["iters"] for r in ones) < 0.8 * sum(r["cg_iters"] for r in ones)


def test_server_cancellation(serve_results):
    rows = [r for r in serve_results if r["case"] == "serve_cancel"]
    assert len(rows) == 6
    r0 = rows[0]
    assert (r0["c_queued"], r0["c_running"], r0["c_again"], r0["c_unknown"]) == (0, 0, 11, 11)
    # request 5, cancelled in the queue: never admitted, no iterations
    assert rows[5]["status"] == 10 and rows[5]["admitted"] == -1.0 and rows[5]["iters"] == 0
    # request 0, cancelled while running: retired with its iterate so far
    assert rows[0]["status"] == 10 and rows[0]["admitted"] >= 0.0 and rows[0]["iters"] > 0
    for r in rows[1:5]:
        assert r["status"] == 0, r
    for r in rows:
        assert r["submitted"] >= 0.0 and r["completed"] >= r["submitted"], r
        if r["admitted"] >= 0.0:
            assert r["submitted"] <= r["admitted"] <= r["completed"], r


def test_server_timeout(serve_results):
    rows = [r for r in serve_results if r["case"] == "serve_timeout"]
    assert len(rows) == 6
    # the limit is set after the first replay: running requests retire with
    # their iterations so far, queued ones without running
    assert all(r["status"] == 10 for r in rows), rows
    assert any(r["admitted"] >= 0.0 and r["iters"] > 0 for r in rows)
    assert any(r["admitted"] == -1.0 and r["iters"] == 0 for r in rows)


def test_server_pcg_refusals(serve_results):
    r = next(r for r in serve_results if r["case"] == "serve_pcg_refusals")
    assert r["create"] == 0 and r["pen"] == 3 and r["busy"] == 1      # UNSUPPORTED, INVALID
