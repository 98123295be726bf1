"""The BDF step controller's failure handling and stop conditions, forced from
the host: tests/ctl/ts_controller_harness.cpp drives the controller's own
host functions (nmag-src_amd/csrc/nsim_tsctl.h, the code the device runs in
its one-thread control kernel) with made-up Newton results.  Every scenario
must end, and end in the documented state.  CPU only."""
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
    out = subprocess.run([exe], check=True, capture_output=True, text=True, timeout=60).stdout
    return {r["name"]: r for r in map(json.loads, out.splitlines())}


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
