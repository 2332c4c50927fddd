"""Error paths of the C ABI under injected device failures, on the CPU
replay build (tests/replay/libintegral_replay.so).  Prints one JSON line
per scenario:

  launch    the row-sum J/K kernel's launches fail: the J/K call must
            report a device error, and the next call, with launches
            working again, must succeed on the same contractor;
  alloc     the first device allocation of a contractor's creation fails:
            creation must report a device error and leave nothing
            allocated; creating it again must succeed;
  replan    the allocation of a re-planned call (a new FP32 cut grows the
            quartet-set tables) fails: the call must report a device error,
            and the same call again must succeed;

with the replay's launch checks staying clean throughout."""
import ctypes
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import integral_amd as ig  # noqa: E402
from integral_amd import synthetic  # noqa: E402
from oracle import oracle as O  # noqa: E402


def main():
    ig.LIB_PATH = os.path.join(HERE, "libintegral_replay.so")
    lib = ig.library()
    lib.replay_set_schwarz.argtypes = [ctypes.c_void_p, ctypes.c_int]
    lib.replay_fail_malloc_after.argtypes = [ctypes.c_longlong]
    lib.replay_fail_launch.argtypes = [ctypes.c_char_p]
    lib.replay_live_allocations.restype = ctypes.c_longlong
    shells, dens, tau = synthetic.direct_jk_config()
    q = np.ascontiguousarray(O.Basis(shells).schwarz(), dtype=np.float64)
    lib.replay_set_schwarz(q.ctypes.data, len(shells))
    basis = synthetic.to_basis(shells)
    dens = list(dens)

    def attempt(fn):
        try:
            fn()
            return None
        except ig.IntegralError as e:
            return "%s: %s" % (type(e).__name__, e)

    # launch failure
    dc = basis.direct_contractor()
    lib.replay_reset()
    lib.replay_fail_launch(b"jk_rowacc_kernel")
    err = attempt(lambda: dc.jk(dens, threshold=tau))
    lib.replay_fail_launch(None)
    again = attempt(lambda: dc.jk(dens, threshold=tau))
    print(json.dumps({"scenario": "launch", "error": err, "next_call_error": again,
                      "violations": lib.replay_violation_count()}), flush=True)

    # allocation failure while a contractor is created
    before = lib.replay_live_allocations()
    lib.replay_fail_malloc_after(0)
    err = attempt(lambda: basis.direct_contractor())
    lib.replay_fail_malloc_after(-1)
    left = lib.replay_live_allocations() - before
    again = attempt(lambda: basis.direct_contractor().jk(dens, threshold=tau))
    print(json.dumps({"scenario": "alloc", "error": err, "left_allocated": int(left),
                      "next_call_error": again,
                      "violations": lib.replay_violation_count()}), flush=True)

    # allocation failure while a call re-plans for a new FP32 cut
    dc.set_mixed_precision(1e-6)
    lib.replay_fail_malloc_after(0)
    err = attempt(lambda: dc.jk(dens, threshold=tau))
    lib.replay_fail_malloc_after(-1)
    again = attempt(lambda: dc.jk(dens, threshold=tau))
    print(json.dumps({"scenario": "replan", "error": err, "next_call_error": again,
                      "violations": lib.replay_violation_count()}), flush=True)


if __name__ == "__main__":
    main()
