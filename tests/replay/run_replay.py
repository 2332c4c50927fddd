"""Run J/K builds of a workload through the CPU replay build of the library
(tests/replay/libintegral_replay.so: the library's host code over a host
stand-in for the HIP runtime that checks every J/K launch) and print one
JSON line per build: launches, quartets checked, violations.

Usage: python tests/replay/run_replay.py WORKLOAD [cut ...]
  WORKLOAD  direct (the nao 520 s,p,d molecule, two densities, tau 1e-12)
            or highl (the nao 565 molecule with g and l = 6 shells, one
            density) or grad (the nao 390 s,p,d molecule: the adjoint
            forces' launches first, then the J/K builds) or eri (the dense
            tensor of the 10-atom nao 50 molecule) or batch (the batched
            plans of six small molecules through ig_batch_check) or
            batchgraph (the same molecules through the batched graph: its
            capture and every graph launch);
  cut       FP32 cut of a mixed build (0: FP64); the default sequence is
            FP64, mixed below 1e-6, mixed below 1e-8, FP64 again, all on one
            contractor (the re-planning between them is part of the check).

With REPLAY_AUTOTUNE set, the contractor's autotuner runs first (every
kernel variant of every class pair at full size) and its launches are
checked as well.  With REPLAY_SCRATCH_PATHS set, the builds also run
through a worker's own workspace and through the no-allocation entry (two
calls, the second on one density).

The Schwarz bounds, which the library computes with a device kernel, come
from the oracle (the replay writes them where that kernel would)."""
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


def batch_graph(lib):
    """The batched graph (compiled in for this build) over the sequence of
    the batched GPU tests: one capture and four graph launches with
    allocations and frees between them; re-captures for a new threshold
    with J only and new output buffers, for FP32 below 1e-8 and back to
    FP64; autotune and a loaded schedule, each followed by a re-capture.
    Each call prints one JSON line.  While a stream captures, launches are
    recorded, not run, with a copy of their arguments; every graph launch
    runs the recorded list against the memory as it is at that launch, so a
    pointer baked into the graph that has gone stale is reported there."""
    os.environ["REPLAY_UNIT_SCHWARZ"] = "1"
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from test_gpu_batch import batch_molecules
    vp = ctypes.c_void_p
    lib.hipMalloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
    lib.hipFree.argtypes = [vp]
    lib.ig_batch_jk_device.argtypes = [vp, vp, ctypes.c_double, ctypes.c_int, vp, vp, vp]
    lib.ig_batch_autotune.argtypes = [vp, ctypes.c_size_t, vp, ctypes.c_double, ctypes.c_int,
                                      ctypes.c_int, vp]
    lib.replay_graph_launches.restype = ctypes.c_longlong
    mols = batch_molecules()
    bases = [synthetic.to_basis(s) for s, _ in mols]

    def dev(a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = vp()
        assert lib.hipMalloc(ctypes.byref(p), a.nbytes) == 0
        ctypes.memmove(p, a.ctypes.data, a.nbytes)
        return p

    def arr(ps):
        return (vp * len(ps))(*[p.value for p in ps])

    def per_molecule(tau):  # the reference builds between the batched calls
        for (_, D), b in zip(mols, bases):
            b.direct_contractor().jk([D], threshold=tau)

    def call(step, bc, dens, J, K, tau):
        lib.replay_reset()
        rc = lib.ig_batch_jk_device(bc._h, arr(dens), tau, 1 if K is not None else 0, arr(J),
                                    arr(K) if K is not None else None, None)
        n = lib.replay_violation_count()
        print(json.dumps({"workload": "batchgraph", "step": step, "rc": rc,
                          "message": lib.ig_last_error_message().decode() if rc else "",
                          "graph_launches": lib.replay_graph_launches(),
                          "jk_launches": lib.replay_jk_launches(),
                          "quartets_checked": lib.replay_quartets(), "violations": n,
                          "first": [lib.replay_violation(k).decode() for k in range(min(n, 5))]}),
              flush=True)

    zeros = [np.zeros_like(D) for _, D in mols]
    bc = ig.BatchContractor(bases)
    per_molecule(1e-12)
    dens = [dev(D) for _, D in mols]
    J = [dev(z) for z in zeros]
    K = [dev(z) for z in zeros]
    for rep in range(4):
        call("replay %d" % rep, bc, dens, J, K, 1e-12)
        tmp = vp()  # other work between the replays
        lib.hipMalloc(ctypes.byref(tmp), 512 * 512 * 8)
        lib.hipFree(tmp)
    per_molecule(1e-9)
    J2 = [dev(z) for z in zeros]
    for rep in range(2):
        call("J only, tau 1e-9, new buffers %d" % rep, bc, dens, J2, None, 1e-9)
    for p in J2:
        lib.hipFree(p)
    bc.set_mixed_precision(1e-8)
    call("FP32 below 1e-8", bc, dens, J, K, 1e-12)
    bc.set_mixed_precision(0.0)
    call("FP64 again", bc, dens, J, K, 1e-12)
    m = int(np.argmax([D.shape[0] for _, D in mols]))
    rc = lib.ig_batch_autotune(bc._h, m, dens[m], 1e-12, 1, 1, None)
    assert rc == 0, lib.ig_last_error_message()
    for rep in range(2):
        call("after autotune %d" % rep, bc, dens, J, K, 1e-12)
    sched = bases[0].direct_contractor().schedule()
    bc.set_schedule([dict(e, variant=0) for e in sched])
    call("loaded schedule", bc, dens, J, K, 1e-12)


def main():
    workload = sys.argv[1] if len(sys.argv) > 1 else "direct"
    cuts = [float(c) for c in sys.argv[2:]] or [0.0, 1e-6, 1e-8, 0.0]
    ig.LIB_PATH = os.path.join(HERE, "libintegral_replay.so")
    lib = ig.library()
    lib.replay_set_schwarz.argtypes = [ctypes.c_void_p, ctypes.c_int]
    lib.replay_violation.restype = ctypes.c_char_p
    lib.replay_violation.argtypes = [ctypes.c_int]
    lib.replay_jk_launches.restype = ctypes.c_longlong
    lib.replay_quartets.restype = ctypes.c_longlong
    lib.replay_malloc_calls.restype = ctypes.c_longlong
    if workload == "batch":
        # the batched plans of the six molecules of the batched tests, with
        # unit Schwarz bounds (every quartet kept), through the library's
        # own host walk of the batched launches (ig_batch_check)
        os.environ["REPLAY_UNIT_SCHWARZ"] = "1"
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from test_gpu_batch import batch_molecules
        mols = batch_molecules()
        bc = ig.BatchContractor([synthetic.to_basis(s) for s, _ in mols])
        for tau, want_k, cut in ((1e-12, True, 0.0), (1e-12, False, 0.0), (1e-9, True, 0.0),
                                 (1e-12, True, 1e-8)):
            bc.set_mixed_precision(cut)
            lib.replay_reset()
            msg = bc.check(tau, want_k)
            print(json.dumps({"workload": workload, "tau": tau, "exchange": want_k,
                              "fp32_below": cut, "check": msg,
                              "violations": lib.replay_violation_count()}), flush=True)
        return
    if workload == "batchgraph":
        batch_graph(lib)
        return
    if workload == "direct":
        shells, dens, tau = synthetic.direct_jk_config()
    elif workload == "eri":
        shells, _ = synthetic.dense_eri_config()
        dens, tau = [], 0.0
    elif workload == "grad":
        shells, D = synthetic.gradient_config()
        dens, tau = [D], 1e-12
    else:
        shells, D = synthetic.high_l_config()
        dens, tau = [D], 1e-12
    q = np.ascontiguousarray(O.Basis(shells).schwarz(), dtype=np.float64)
    lib.replay_set_schwarz(q.ctypes.data, len(shells))
    basis = synthetic.to_basis(shells)
    if workload == "eri":
        # the dense nao^4 tensor (every canonical quartet, both engines'
        # kernels as the library picks them)
        lib.replay_reset()
        basis.eri()
        n = lib.replay_violation_count()
        print(json.dumps({"workload": workload, "dense": True,
                          "jk_launches": lib.replay_jk_launches(),
                          "quartets_checked": lib.replay_quartets(), "violations": n,
                          "first": [lib.replay_violation(k).decode() for k in range(min(n, 5))]}),
              flush=True)
        return
    if workload == "grad":
        # the J/K energy's contraction and the adjoint forces (Coulomb 1,
        # exchange -1/2), as one geometry of the energy-and-forces workload
        gc = basis.direct_grad_contractor()
        lib.replay_reset()
        dg = gc.density([ig.GradDensityTerm(0.5, D, D)], [ig.GradDensityTerm(-0.25, D, D)])
        gc.contract(dg, threshold=tau)
        n = lib.replay_violation_count()
        print(json.dumps({"workload": workload, "forces": True,
                          "jk_launches": lib.replay_jk_launches(),
                          "quartets_checked": lib.replay_quartets(), "violations": n,
                          "first": [lib.replay_violation(k).decode() for k in range(min(n, 5))]}),
              flush=True)
    dc = basis.direct_contractor()
    if os.environ.get("REPLAY_AUTOTUNE"):
        # the autotuner launches every kernel variant of every class pair at
        # full size; its "device" densities are an allocation of the stand-in
        lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        lib.ig_direct_autotune.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                           ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                           ctypes.c_void_p]
        d = np.ascontiguousarray(np.stack(dens), dtype=np.float64)
        dptr = ctypes.c_void_p()
        lib.hipMalloc(ctypes.byref(dptr), d.nbytes)
        ctypes.memmove(dptr, d.ctypes.data, d.nbytes)
        lib.replay_reset()
        rc = lib.ig_direct_autotune(dc._h, dptr, len(dens), float(tau), 1, 1, None)
        n = lib.replay_violation_count()
        print(json.dumps({"workload": workload, "autotune_rc": rc,
                          "jk_launches": lib.replay_jk_launches(),
                          "quartets_checked": lib.replay_quartets(), "violations": n,
                          "first": [lib.replay_violation(k).decode() for k in range(min(n, 5))]}),
              flush=True)
    def report(**kw):
        n = lib.replay_violation_count()
        rec = {"workload": workload}
        rec.update(kw)
        rec.update({"jk_launches": lib.replay_jk_launches(),
                    "quartets_checked": lib.replay_quartets(), "violations": n,
                    "first": [lib.replay_violation(k).decode() for k in range(min(n, 5))]})
        print(json.dumps(rec), flush=True)

    for cut in cuts:
        lib.replay_reset()
        dc.set_mixed_precision(cut)
        dc.jk(list(dens), threshold=tau)
        report(fp32_below=cut)
    if os.environ.get("REPLAY_SCRATCH_PATHS"):
        # a worker's own workspace (its scratch, the contractor's tables)
        ws = ig.DirectWorkspace()
        for cut in (0.0, 1e-6):
            lib.replay_reset()
            dc.set_mixed_precision(cut)
            dc.jk(list(dens), threshold=tau, workspace=ws)
            report(path="workspace", fp32_below=cut)
        # the no-allocation entry: every view carved from the fixed device
        # array; two calls, the second on fewer densities
        dc.set_mixed_precision(0.0)
        lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        lib.ig_direct_jk_device_noalloc.argtypes = [
            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_double, ctypes.c_int,
            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
        n = basis.nao()
        for nd in (len(dens), 1):
            d = np.ascontiguousarray(np.stack(dens[:nd]), dtype=np.float64)
            bufs = []
            for _ in range(3):
                p = ctypes.c_void_p()
                lib.hipMalloc(ctypes.byref(p), d.nbytes)
                bufs.append(p)
            ctypes.memmove(bufs[0], d.ctypes.data, d.nbytes)
            lib.replay_reset()
            rc = lib.ig_direct_jk_device_noalloc(dc._h, bufs[0], nd, float(tau), 1, bufs[1],
                                                 bufs[2], 1, None)
            msg = lib.ig_last_error_message().decode() if rc else ""
            report(path="noalloc", ndens=nd, rc=rc, message=msg,
                   device_allocations=lib.replay_malloc_calls())


if __name__ == "__main__":
    main()
