"""CPU replay of the direct J/K launches (no GPU): the library's host code,
compiled host-only over a host stand-in for the HIP runtime
(tests/replay/hip_replay.cpp), plans and launches the nao 520 s,p,d
workload's builds (FP64 and mixed at 1e-6 and 1e-8 on one contractor, and
the autotuner's trials), the nao 565 high-l workload's and the nao 390
workload's forces, and every J/K and gradient launch is checked: its views
against the live allocations, the kernel's LDS and scratch, and every
quartet of its set against the pair, shell, function and row-sum tables.
The planted-mistake run asks the checker for one row pointer more than each
set holds and must be reported."""
import json
import os
import shutil
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
REPLAY = os.path.join(HERE, "replay")


@pytest.fixture(scope="module")
def replay_lib():
    if not shutil.which("/opt/rocm/bin/hipcc"):
        pytest.skip("hipcc not present: the replay build compiles the library's host code")
    subprocess.run(["make", "-j8", "-C", REPLAY], check=True, capture_output=True, text=True,
                   timeout=900)
    return os.path.join(REPLAY, "libintegral_replay.so")


def run(workload, cuts, env=None):
    e = dict(os.environ)
    e.update(env or {})
    out = subprocess.run([sys.executable, os.path.join(REPLAY, "run_replay.py"), workload]
                         + [str(c) for c in cuts], check=True, capture_output=True, text=True,
                         timeout=900, env=e)
    return [json.loads(l) for l in out.stdout.splitlines() if l.startswith("{")]


def test_direct_jk_launches_stay_in_range_fp64_and_mixed(replay_lib):
    recs = run("direct", [0.0, 1e-6, 1e-8, 0.0])
    assert [r["fp32_below"] for r in recs] == [0.0, 1e-6, 1e-8, 0.0]
    for r in recs:
        assert r["violations"] == 0, r["first"]
        assert r["jk_launches"] > 0 and r["quartets_checked"] > 8e7
    # the mixed builds launch the FP64 and FP32 part of each class pair
    assert recs[1]["jk_launches"] > recs[0]["jk_launches"]
    assert recs[3] == recs[0]


def test_high_l_jk_launches_stay_in_range_fp64_and_mixed(replay_lib):
    # the nao 565 workload with g and l = 6 shells: its compiled classes and
    # the runtime-l per-quartet and sliced kernels (offset table, block
    # slots, root-batch and digest LDS)
    recs = run("highl", [0.0, 1e-7])
    for r in recs:
        assert r["violations"] == 0, r["first"]
        assert r["jk_launches"] > 200 and r["quartets_checked"] > 5e6


def test_gradient_launches_stay_in_range(replay_lib):
    # the nao 390 energy-and-forces workload: adjoint force kernels, then J/K
    recs = run("grad", [0.0])
    assert recs[0].get("forces") and recs[0]["jk_launches"] > 50
    for r in recs:
        assert r["violations"] == 0, r["first"]
        assert r["quartets_checked"] > 1e7


def test_workspace_and_no_allocation_paths_stay_in_range(replay_lib):
    recs = run("direct", [0.0], env={"REPLAY_SCRATCH_PATHS": "1"})
    ws = [r for r in recs if r.get("path") == "workspace"]
    na = [r for r in recs if r.get("path") == "noalloc"]
    assert [r["fp32_below"] for r in ws] == [0.0, 1e-6]
    assert [r["ndens"] for r in na] == [2, 1]
    for r in recs:
        assert r["violations"] == 0, r["first"]
        assert r["jk_launches"] > 0
    for r in na:  # every view carved from the fixed array: no allocation
        assert r["rc"] == 0, r["message"]
        assert r["device_allocations"] == 0


def test_dense_eri_launches_stay_in_range(replay_lib):
    rec = run("eri", [])[0]
    assert rec["dense"] and rec["jk_launches"] > 0 and rec["quartets_checked"] > 1e5
    assert rec["violations"] == 0, rec["first"]


def test_batched_plans_pass_the_library_host_walk(replay_lib):
    # ig_batch_check on the CPU: the six molecules of the batched tests,
    # every quartet kept (unit Schwarz bounds), FP64 and FP32 below 1e-8
    recs = run("batch", [])
    assert len(recs) == 4
    for r in recs:
        assert r["check"] == "" and r["violations"] == 0, r


def test_batched_graph_launches_stay_in_range_over_replays_and_recaptures(replay_lib):
    # the batched graph path (compiled into this build only) over the batched
    # GPU tests' sequence: every launch recorded at capture is run again, with
    # its baked arguments, at each graph launch against the live allocations;
    # no capture-unsafe call and no launch off the capturing stream
    recs = run("batchgraph", [])
    assert [r["graph_launches"] for r in recs] == list(range(1, 12))
    for r in recs:
        assert r["rc"] == 0, r["message"]
        assert r["violations"] == 0, (r["step"], r["first"])
        assert r["jk_launches"] > 100 and r["quartets_checked"] > 5e5
    # replays of one graph check the same launches
    assert len({(r["jk_launches"], r["quartets_checked"]) for r in recs[:4]}) == 1


def test_autotuner_trial_launches_stay_in_range(replay_lib):
    # every kernel variant of every class pair at full size
    recs = run("direct", [0.0], env={"REPLAY_AUTOTUNE": "1"})
    tune = recs[0]
    assert tune["autotune_rc"] == 0
    assert tune["violations"] == 0, tune["first"]
    assert tune["jk_launches"] > 100 and tune["quartets_checked"] > 3e8
    assert recs[1]["violations"] == 0, recs[1]["first"]


def test_replay_reports_a_planted_undersized_row_pointer_array(replay_lib):
    recs = run("direct", [0.0], env={"REPLAY_PLANT": "rowptr"})
    assert recs[0]["violations"] > 0
    assert "row pointers" in recs[0]["first"][0]


def test_injected_device_failures_are_reported_and_survived(replay_lib):
    """A failed kernel launch, a failed allocation while a contractor is
    created, and a failed allocation while a call re-plans for a new FP32
    cut: each comes back as a DeviceError naming the step, leaves nothing
    allocated behind, and the next call on the same objects succeeds."""
    out = subprocess.run([sys.executable, os.path.join(REPLAY, "run_failures.py")],
                         check=True, capture_output=True, text=True, timeout=900)
    recs = {r["scenario"]: r for r in (json.loads(l) for l in out.stdout.splitlines()
                                       if l.startswith("{"))}
    assert set(recs) == {"launch", "alloc", "replan"}
    assert recs["launch"]["error"].startswith("DeviceError: jk launch")
    assert recs["alloc"]["error"].startswith("DeviceError: hipMalloc")
    assert recs["replan"]["error"].startswith("DeviceError: hipMalloc")
    assert recs["alloc"]["left_allocated"] == 0
    for r in recs.values():
        assert r["next_call_error"] is None
        assert r["violations"] == 0
