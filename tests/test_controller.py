"""The BDF step controller, driven from the host: tests/ctl/ts_controller_harness.cpp
runs the controller's own host functions (nmag-src_amd/csrc/nsim_tsctl.h, the
code the device runs at the end of every Newton iteration) through forced
failure and stop scenarios with made-up Newton results -- every scenario must
end, and end in the documented state -- and through whole integrations of a
linear ODE with exact Newton solves, which pin the BDF orders 1 to 5 and the
order selection against the exact solution.  CPU only."""
import json
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

ERROR, DONE = 3, 2       # TsMode kTsError, kTsDone
MAX_ATTEMPTS, MAX_COR = 20, 3


@pytest.fixture(scope="module")
def runs(tmp_path_factory):
    if not (os.path.exists(HIPCC) or shutil.which("hipcc")):
        pytest.skip("hipcc is needed to build the controller harness")
    exe = str(tmp_path_factory.mktemp("ctl") / "harness")
    subprocess.run([HIPCC if os.path.exists(HIPCC) else "hipcc", "-O1", "-std=c++17",
                    "--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang",
                    "-packed-fp32-ops", "-I" + os.path.join(ROOT, "nmag-src_amd", "csrc"),
                    os.path.join(ROOT, "tests", "ctl", "ts_controller_harness.cpp"),
                    "-o", exe], check=True, capture_output=True, timeout=300)
    out = subprocess.run([exe], check=True, capture_output=True, text=True, timeout=120).stdout
    rows = [json.loads(line) for line in out.splitlines()]
    res = {r["name"]: r for r in rows}
    res["_all"] = rows
    return res


def test_error_test_failures_end_in_an_error(runs):
    r = runs["error_test_always_fails"]
    # each rejection takes h to 0.2 h; the step-size floor 1e-14 ends it
    assert r["mode"] == ERROR and r["err_code"] == 2 and r["halt"] == 1
    assert r["steps"] == 0 and r["err_fails"] == r["cycles"] <= MAX_ATTEMPTS


def test_newton_divergence_ends_after_bounded_attempts(runs):
    r = runs["newton_diverges"]
    assert r["mode"] == ERROR and r["err_code"] == 2
    assert r["conv_fails"] == MAX_ATTEMPTS and r["cycles"] == MAX_ATTEMPTS * MAX_COR
    n = runs["nan_correction"]
    assert n["mode"] == ERROR and n["conv_fails"] == MAX_ATTEMPTS and n["cycles"] == MAX_ATTEMPTS


def test_stops_at_t_final_and_at_the_step_budget(runs):
    r = runs["smooth_exact_stop"]
    assert r["mode"] == DONE and r["err_code"] == 0 and r["tn"] == r["t_final"] == 1.0
    b = runs["smooth_max_steps"]
    assert b["mode"] == DONE and b["steps_call"] == 7 and b["tn"] < b["t_final"]


def test_two_rejections_drop_the_order(runs):
    r = runs["two_rejections_then_accept"]
    assert r["mode"] == DONE and r["err_fails"] == 2 and r["steps"] == 1
    # error-test rejections at err >> 1 take h to 0.2 h each
    assert r["h"] == pytest.approx(0.01 * 0.2 * 0.2, rel=1e-12)
    assert r["q_last"] == 1


# Whole integrations of y' = A y (a damped rotation and a decay, exact
# solution known) to t = 10 by the controller with exact Newton solves on the
# host: BDF of fixed order q, and with order selection up to max_order.


def _integrations(runs_all, name):
    return {(r["max_order"], r["tol"]): r for r in runs_all if r["name"] == name}


@pytest.fixture(scope="module")
def runs_all(runs):
    return runs["_all"]


@pytest.mark.parametrize("q", [1, 2, 3, 4, 5])
def test_fixed_order_step_counts_and_errors_follow_the_order(runs_all, q):
    """Local-error control at order q takes h ~ tol^(1/(q+1)): a 100-fold
    smaller tolerance multiplies the steps by 100^(1/(q+1)) and divides the
    global error by 100^(q/(q+1))."""
    fx = _integrations(runs_all, "fixed")
    for lo, hi in ((1e-6, 1e-8), (1e-8, 1e-10)):
        a, b = fx[(q, lo)], fx[(q, hi)]
        assert a["mode"] == DONE == b["mode"] and a["tn"] == 10.0
        assert a["order_steps"][q - 1] >= 0.9 * a["steps"]
        steps = b["steps"] / a["steps"]
        assert 0.8 * 100 ** (1 / (q + 1)) < steps < 1.25 * 100 ** (1 / (q + 1)), (q, steps)
        gain = a["err"] / b["err"]
        assert 0.6 * 100 ** (q / (q + 1)) < gain < 1.6 * 100 ** (q / (q + 1)), (q, gain)


@pytest.mark.parametrize("q", [3, 4, 5])
def test_order_selection_climbs_to_the_maximum_order(runs_all, q):
    sel, fx = _integrations(runs_all, "select"), _integrations(runs_all, "fixed")
    for tol in (1e-6, 1e-8, 1e-10):
        r = sel[(q, tol)]
        assert r["mode"] == DONE and r["q_last"] == q
        assert r["order_steps"][q - 1] > 0.7 * r["steps"]
        # as accurate as the fixed order and no more steps
        assert r["err"] < 1.5 * fx[(q, tol)]["err"]
        assert r["steps"] <= fx[(q, tol)]["steps"]
    # order 5 against order 2 at tight tolerance: a fraction of the steps
    assert sel[(5, 1e-10)]["steps"] < fx[(2, 1e-10)]["steps"] / 20
