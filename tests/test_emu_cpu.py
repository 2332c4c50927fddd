"""The library's kernels executed on the host, no GPU (tests/emu): the kernel
sources compiled as C++ against a shim of the HIP device API, each workgroup
run with its lanes as fibers of one host thread (wave operations resolved
over the wavefront's active lanes, barriers over the workgroup's live
lanes), with host AddressSanitizer on every global and LDS access, the host
code and the runtime stand-in of tests/replay around them.  Results are
compared with the oracle at the GPU tests' tolerances: direct J/K of the
batched tests' six molecules (FP64 and FP32 below 1e-6), the batched graph
over capture, replay and re-captures (also under a random wavefront
schedule), direct J/K, dense ERI and gradients with g, h and i shells (the
runtime-l kernels) and the s..f gradients."""
import json
import os
import shutil
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
EMU = os.path.join(HERE, "emu")


@pytest.fixture(scope="module")
def emu_env():
    if not shutil.which("/opt/rocm/bin/hipcc") or not shutil.which("g++"):
        pytest.skip("hipcc and g++ build the host-executed library")
    asan = subprocess.run(["g++", "-print-file-name=libasan.so"], capture_output=True,
                          text=True).stdout.strip()
    if not os.path.isabs(asan) or not os.path.exists(asan):
        pytest.skip("the host AddressSanitizer runtime is not installed")
    subprocess.run(["make", "-j8", "-C", EMU], check=True, capture_output=True, text=True,
                   timeout=1200)
    env = dict(os.environ)
    env.update({"LD_PRELOAD": asan, "ASAN_OPTIONS": "detect_leaks=0", "REPLAY_CU_COUNT": "2"})
    return env


def run(env, workload):
    out = subprocess.run([sys.executable, os.path.join(EMU, "run_emu.py"), workload],
                         capture_output=True, text=True, timeout=1200, env=env)
    assert out.returncode == 0, out.stderr[-3000:]
    return [json.loads(l) for l in out.stdout.splitlines() if l.startswith("{")]


def test_direct_jk_on_host_executed_kernels_matches_oracle(emu_env):
    recs = run(emu_env, "direct")
    assert len(recs) == 12
    for r in recs:
        assert r["violations"] == 0, r["first"]
        bar = 1e-12 if r["fp32_below"] == 0.0 else 1e-9
        assert r["dJ"] < bar and r["dK"] < bar, r


def test_batched_graph_on_host_executed_kernels_matches_oracle(emu_env):
    # the batched graph's own kernel (work-queue chunks over the six
    # molecules, nao 1-126) over graph launches and re-captures
    recs = run(emu_env, "batchgraph")
    assert [r["graph_launches"] for r in recs] == [1, 2, 3, 4, 5, 6, 7]
    for r in recs:
        assert r["rc"] == 0 and r["violations"] == 0, r
        bar = 1e-9 if "FP32" in r["step"] else 1e-12
        assert r["dJ"] < bar and r["dK"] < bar, r


def test_dense_eri_on_host_executed_kernels_matches_oracle(emu_env):
    recs = run(emu_env, "eri")
    assert [r["basis"] for r in recs] == ["s..f", "l=4", "l=5", "l=6"]
    for r in recs:
        assert r["violations"] == 0, r["first"]
        assert r["dERI"] < 1e-12 * max(1.0, r["scale"]), r


def test_gradients_on_host_executed_kernels_match_oracle(emu_env):
    recs = run(emu_env, "grad")
    assert [r["basis"] for r in recs] == ["s..f cart", "s..f sph", "l=4", "l=5"]
    for r in recs:
        assert r["violations"] == 0, r["first"]
        assert r["dF"] < 1e-12, r


def test_runtime_l_jk_on_host_executed_kernels_matches_oracle(emu_env):
    # g, h and i shells: the per-quartet runtime-l kernel, and the sliced one
    # for the (ii|ii) block (614656 elements, above the 262144 split)
    recs = run(emu_env, "highl")
    assert len(recs) == 12
    for r in recs:
        assert r["violations"] == 0, r["first"]
        bar = 1e-12 if r["fp32_below"] == 0.0 else 1e-9
        assert r["dJ"] < bar and r["dK"] < bar, r


def test_batched_graph_under_a_random_wavefront_schedule(emu_env):
    # wavefronts advanced one at a time in a seeded random order, so they
    # drift apart between barriers as on the hardware: the LDS row sums,
    # flushes and work-queue reads still give the oracle's J and K
    env = dict(emu_env, EMU_SCHEDULE="random:7", EMU_BATCH_SHORT="1")
    recs = run(env, "batchgraph")
    assert [r["graph_launches"] for r in recs] == [1, 2]
    for r in recs:
        assert r["rc"] == 0 and r["violations"] == 0, r
        bar = 1e-9 if "FP32" in r["step"] else 1e-12
        assert r["dJ"] < bar and r["dK"] < bar, r


def test_workspace_and_no_allocation_entry_on_host_executed_kernels(emu_env):
    # the no-allocation entry twice per molecule (two densities, then one)
    # and on a second molecule after the first: every view carved from the
    # library's fixed array, J and K at the oracle's values, no allocation
    recs = run(emu_env, "scratch")
    assert [(r["path"], r["molecule"]) for r in recs] == [
        ("workspace", 0), ("noalloc", 0), ("noalloc", 0),
        ("workspace", 5), ("noalloc", 5), ("noalloc", 5)]
    for r in recs:
        assert r["violations"] == 0, r["first"]
        assert r["dJ"] < 1e-12 and r["dK"] < 1e-12, r
        if r["path"] == "noalloc":
            assert r["rc"] == 0 and r["device_allocations"] == 0, r


def test_gpu_tests_of_the_next_rows_pass_on_host_executed_kernels(emu_env):
    # the GPU test files of density fitting, the packed ERI and EriBuilder,
    # and the erf operator (closed forms, the omega -> infinity limit, the
    # builder's ranges) run unchanged against the host-executed library
    out = subprocess.run([sys.executable, os.path.join(EMU, "run_gpu_tests.py"),
                          "tests/test_gpu_df.py", "tests/test_gpu_packed.py",
                          "tests/test_gpu_erf.py"], capture_output=True, text=True,
                         timeout=1500, env=emu_env)
    assert out.returncode == 0, out.stdout[-3000:] + out.stderr[-2000:]
    assert "14 passed" in out.stdout and "failed" not in out.stdout, out.stdout[-2000:]
