"""Run the library with its kernels executed on the host
(tests/emu/libintegral_emu.so) and compare with the oracle.  Prints one JSON
line per case: max |difference| from the oracle, and the violations the
runtime stand-in and the lane scheduler recorded (AddressSanitizer ends the
process on an access outside any allocation or outside a workgroup's LDS).

Usage: python tests/emu/run_emu.py WORKLOAD  (run with the sanitizer's
runtime preloaded: LD_PRELOAD=libasan.so, ASAN_OPTIONS=detect_leaks=0)
  direct      the six molecules of the batched GPU tests (nao 1-126, s..f,
              Cartesian and spherical), direct J/K at tau 1e-12 in FP64 and
              with FP32 below 1e-6;
  batchgraph  the same molecules through the batched graph: a capture, a
              second graph launch, re-captures for J only at tau 1e-9, for
              FP32 below 1e-8 and back to FP64, after autotune and with
              every class on the team kernels; every molecule's J and K
              after each graph launch (EMU_BATCH_SHORT: the first two
              graph launches only);
  eri         dense ERI tensors: the s..f reference basis and three-shell
              bases with a g, h and i shell (the runtime-l kernels);
  grad        direct gradients: the s..f reference basis (Cartesian and
              spherical) and three-shell bases with a g and an h shell;
  highl       direct J/K of three-shell bases with a g, h and i shell (the
              runtime-l per-quartet kernel, and the sliced one for the
              (ii|ii) block of 614656 elements), FP64 and FP32 below 1e-7;
  scratch     J/K through a worker's workspace and through the
              no-allocation entry (two densities, then one).
"""
import ctypes
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import integral_amd as ig  # noqa: E402
from integral_amd import synthetic  # noqa: E402
from oracle import oracle as O  # noqa: E402

ig.LIB_PATH = os.path.join(HERE, "libintegral_emu.so")
lib = ig.library()
lib.replay_violation.restype = ctypes.c_char_p
lib.replay_violation.argtypes = [ctypes.c_int]
lib.replay_graph_launches.restype = ctypes.c_longlong


def emit(**kw):
    n = lib.replay_violation_count()
    kw.update({"violations": n,
               "first": [lib.replay_violation(k).decode() for k in range(min(n, 5))]})
    print(json.dumps(kw), flush=True)


def maxdiff(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a.reshape(b.shape) - b))) if b.size else 0.0


def direct():
    from test_gpu_batch import batch_molecules
    for m, (shells, D) in enumerate(batch_molecules()):
        ob = O.Basis(shells)
        oJ, oK = ob.direct_jk([D], 1e-12)
        dc = synthetic.to_basis(shells).direct_contractor()
        for cut in (0.0, 1e-6):
            dc.set_mixed_precision(cut)
            lib.replay_reset()
            t = time.time()
            J, K = dc.jk([D], threshold=1e-12)
            emit(workload="direct", molecule=m, nao=D.shape[0], fp32_below=cut,
                 seconds=round(time.time() - t, 2), dJ=maxdiff(J[0], oJ[0]),
                 dK=maxdiff(K[0], oK[0]))


def batch_graph():
    from test_gpu_batch import batch_molecules
    vp = ctypes.c_void_p
    lib.hipMalloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
    lib.ig_batch_jk_device.argtypes = [vp, vp, ctypes.c_double, ctypes.c_int, vp, vp, vp]
    mols = batch_molecules()
    ref = {tau: [O.Basis(s).direct_jk([D], tau) for s, D in mols] for tau in (1e-12, 1e-9)}

    def dev(a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = vp()
        assert lib.hipMalloc(ctypes.byref(p), a.nbytes) == 0
        ctypes.memmove(p, a.ctypes.data, a.nbytes)
        return p

    def host(p, n):
        out = np.empty((n, n))
        ctypes.memmove(out.ctypes.data, p, out.nbytes)
        return out

    def arr(ps):
        return (vp * len(ps))(*[p.value for p in ps])

    def call(step, bc, dens, J, K, tau):
        for ps in (J, K) if K is not None else (J,):
            for p, (_, D) in zip(ps, mols):
                z = np.zeros_like(D)
                ctypes.memmove(p, z.ctypes.data, z.nbytes)
        lib.replay_reset()
        t = time.time()
        rc = lib.ig_batch_jk_device(bc._h, arr(dens), tau, 1 if K is not None else 0, arr(J),
                                    arr(K) if K is not None else None, None)
        dj = dk = 0.0
        if rc == 0:
            for m, (_, D) in enumerate(mols):
                n = D.shape[0]
                oJ, oK = ref[tau][m]
                dj = max(dj, maxdiff(host(J[m], n), oJ[0]))
                if K is not None:
                    dk = max(dk, maxdiff(host(K[m], n), oK[0]))
        emit(workload="batchgraph", step=step, rc=rc, seconds=round(time.time() - t, 2),
             graph_launches=lib.replay_graph_launches(), dJ=dj, dK=dk)

    bc = ig.BatchContractor([synthetic.to_basis(s) for s, _ in mols])
    dens = [dev(D) for _, D in mols]
    J = [dev(np.zeros_like(D)) for _, D in mols]
    K = [dev(np.zeros_like(D)) for _, D in mols]
    for rep in range(2):
        call("replay %d" % rep, bc, dens, J, K, 1e-12)
    if os.environ.get("EMU_BATCH_SHORT"):  # the capture and one replay only
        return
    J2 = [dev(np.zeros_like(D)) for _, D in mols]
    call("J only, tau 1e-9", bc, dens, J2, None, 1e-9)
    bc.set_mixed_precision(1e-8)
    call("FP32 below 1e-8", bc, dens, J, K, 1e-12)
    bc.set_mixed_precision(0.0)
    call("FP64 again", bc, dens, J, K, 1e-12)
    # the tuned schedule, then every class on the team kernels (per-molecule
    # launches inside the graph), each captured again
    lib.ig_batch_autotune.argtypes = [vp, ctypes.c_size_t, vp, ctypes.c_double, ctypes.c_int,
                                      ctypes.c_int, vp]
    big = int(np.argmax([D.shape[0] for _, D in mols]))
    assert lib.ig_batch_autotune(bc._h, big, dens[big], 1e-12, 1, 1, None) == 0
    call("after autotune", bc, dens, J, K, 1e-12)
    sched = synthetic.to_basis(mols[0][0]).direct_contractor().schedule()
    bc.set_schedule([dict(e, variant=0) for e in sched])
    call("team kernels schedule", bc, dens, J, K, 1e-12)


def three_shell(l, sph):
    return [(0, (0.0, 0.0, 0.0), [1.3, 0.4], [0.6, 0.5], sph),
            (l, (0.3, -0.2, 0.4), [0.7], [1.0], sph),
            (1, (-0.5, 0.6, -0.1), [0.9], [1.0], sph)]


def eri():
    import bases as B
    cases = [("s..f", B.direct_contractor_basis(True))] + \
            [("l=%d" % l, three_shell(l, True)) for l in (4, 5, 6)]
    for name, shells in cases:
        lib.replay_reset()
        t = time.time()
        g = synthetic.to_basis(shells).eri()
        o = O.Basis(shells).eri()
        emit(workload="eri", basis=name, seconds=round(time.time() - t, 2), dERI=maxdiff(g, o),
             scale=float(np.max(np.abs(o))))


def grad():
    import bases as B
    cases = [("s..f cart", B.grad_basis(False)), ("s..f sph", B.grad_basis(True)),
             ("l=4", three_shell(4, True)), ("l=5", three_shell(5, True))]
    for name, shells in cases:
        n = B.nfunc(shells)
        coul, exch = B.grad_terms(n)
        lib.replay_reset()
        t = time.time()
        gc = synthetic.to_basis(shells).direct_grad_contractor()
        dens = gc.density([ig.GradDensityTerm(*c) for c in coul],
                          [ig.GradDensityTerm(*c) for c in exch])
        g = gc.contract(dens)
        o = O.Basis(shells).direct_grad(coul, exch, 0.0)
        emit(workload="grad", basis=name, seconds=round(time.time() - t, 2), dF=maxdiff(g, o),
             scale=float(np.max(np.abs(o))))


def highl():
    rng = np.random.default_rng(20261203)
    for l in (4, 5, 6):
        for sph in (True, False):
            shells = three_shell(l, sph)
            b = synthetic.to_basis(shells)
            n = b.nao()
            C = rng.normal(size=(n, 2)) * 0.1
            D = C @ C.T
            oJ, oK = O.Basis(shells).direct_jk([D], 1e-12)
            dc = b.direct_contractor()
            for cut in (0.0, 1e-7):
                dc.set_mixed_precision(cut)
                lib.replay_reset()
                t = time.time()
                J, K = dc.jk([D], threshold=1e-12)
                emit(workload="highl", l=l, sph=sph, fp32_below=cut,
                     seconds=round(time.time() - t, 2), dJ=maxdiff(J[0], oJ[0]),
                     dK=maxdiff(K[0], oK[0]), scale=float(np.max(np.abs(oJ[0]))))


def scratch():
    """The J/K builds through a worker's own workspace and through the
    no-allocation entry (views carved from the library's fixed device
    array): two molecules, two densities and then one, against the oracle."""
    from test_gpu_batch import batch_molecules
    vp = ctypes.c_void_p
    lib.hipMalloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
    lib.ig_direct_jk_device_noalloc.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_double,
                                                ctypes.c_int, vp, vp, ctypes.c_int, vp]
    lib.replay_malloc_calls.restype = ctypes.c_longlong
    rng = np.random.default_rng(20261204)
    for m in (0, 5):
        shells, D = batch_molecules()[m]
        n = D.shape[0]
        C = rng.normal(size=(n, 3)) * 0.1
        dens = [D, C @ C.T]
        oJ, oK = O.Basis(shells).direct_jk(dens, 1e-12)
        dc = synthetic.to_basis(shells).direct_contractor()
        ws = ig.DirectWorkspace()
        lib.replay_reset()
        J, K = dc.jk(list(dens), threshold=1e-12, workspace=ws)
        emit(workload="scratch", path="workspace", molecule=m,
             dJ=max(maxdiff(J[k], oJ[k]) for k in range(2)),
             dK=max(maxdiff(K[k], oK[k]) for k in range(2)))
        for nd in (2, 1):
            d = np.ascontiguousarray(np.stack(dens[:nd]), dtype=np.float64)
            bufs = []
            for k in range(3):
                p = vp()
                assert lib.hipMalloc(ctypes.byref(p), d.nbytes) == 0
                z = d if k == 0 else np.zeros_like(d)
                ctypes.memmove(p, z.ctypes.data, z.nbytes)
                bufs.append(p)
            lib.replay_reset()
            rc = lib.ig_direct_jk_device_noalloc(dc._h, bufs[0], nd, 1e-12, 1, bufs[1], bufs[2],
                                                 1, None)
            allocs = lib.replay_malloc_calls()
            out = [np.empty_like(d), np.empty_like(d)]
            for k in range(2):
                ctypes.memmove(out[k].ctypes.data, bufs[k + 1], d.nbytes)
            emit(workload="scratch", path="noalloc", molecule=m, ndens=nd, rc=rc,
                 device_allocations=allocs,
                 dJ=max(maxdiff(out[0][k], oJ[k]) for k in range(nd)),
                 dK=max(maxdiff(out[1][k], oK[k]) for k in range(nd)))


if __name__ == "__main__":
    {"direct": direct, "batchgraph": batch_graph, "eri": eri, "grad": grad,
     "highl": highl, "scratch": scratch}[sys.argv[1]]()
