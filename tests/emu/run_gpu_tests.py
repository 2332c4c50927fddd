"""Run GPU test files whose cases use only the library's host API (no torch
device tensors) against the host-executed kernels: the package is pointed at
tests/emu/libintegral_emu.so before pytest imports the tests.  Run with the
sanitizer's runtime preloaded (see run_emu.py).

Usage: python tests/emu/run_gpu_tests.py TESTFILE... [pytest options]"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import pytest  # noqa: E402

import integral_amd as ig  # noqa: E402

ig.LIB_PATH = os.path.join(HERE, "libintegral_emu.so")

if __name__ == "__main__":
    os.chdir(ROOT)
    sys.exit(pytest.main(["-m", "gpu", "-p", "no:cacheprovider", "-q", "-rs"] + sys.argv[1:]))
