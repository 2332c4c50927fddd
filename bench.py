#!/usr/bin/env python3
"""Benchmark of the integral-direct J/K build on one MI355X.

Workload: 40 atoms in a 14 bohr cube (min separation 1.5 bohr), spherical
s,s,p,p,d shells per atom (s/p 3 primitives, d 1), exponents log-uniform in
[0.08, 40]; nao = 520; two densities D = C C^T with C nao x 20 ~ N(0, 0.1);
Schwarz threshold 1e-12; FP64 throughout.  One step = one full J/K build
(both densities, J and K) from device-resident densities into device J/K.

Prints one JSON line: quartets/s (canonical shell quartets accepted by the
screening, divided by J/K build seconds), the roofline record of the dominant
quartet-class kernel (algorithmic flops / HIP-event time vs the 78.6 TFLOP/s
FP64 peak) and the CPU baseline (the oracle's C restatement of the reference
algorithm, OpenMP over bra pairs, timed on a bounded bra-pair sample).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6
METRIC = "J/K build s & quartets/s, mixed vs FP64 GPU at |ΔJ,K|≤1e-9; grad s/geometry"


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def class_flops(rec, spherical=True, ndens=2):
    from integral_amd import flops
    q, kp = rec["quartets"], rec["prim_quartets"]
    if q == 0:
        return 0.0
    K = kp / q  # uniform per class in this workload
    ls = rec["cls"]
    per = flops.quartet_flops(ls, K, [spherical] * 4, ndens=ndens, c2s_in_loop=False)
    return per * q


def kernel_name(cls, variant):
    """Kernel(s) of one class launch: the team kernel, or the register kernel
    in as many element slices as the class needs (launched back to back)."""
    ncart = [(l + 1) * (l + 2) // 2 for l in cls]
    n4 = ncart[0] * ncart[1] * ncart[2] * ncart[3]
    c = "%d,%d,%d,%d" % tuple(cls)
    if variant == 0:
        return "jk_kernel<Cls<%s>>" % c
    slices = (n4 + 80) // 81
    return "jk_reg_kernel<Cls<%s>, LO, HI> (%d element slice%s)" % (c, slices,
                                                                     "s" if slices > 1 else "")


def cpu_baseline(shells, dens, tau, seconds):
    """Oracle on the host cores over a contiguous sample of canonical bra pairs
    from the middle of the list, grown until it takes about `seconds`."""
    from oracle import oracle as O
    threads = min(16, os.cpu_count() or 1)
    ob = O.Basis(shells)
    q = ob.schwarz(threads=threads)
    bounds = ob.density_shell_bounds(dens)
    npairs = ob.nsh * (ob.nsh + 1) // 2
    b0 = npairs // 2
    m = 4
    while True:
        t0 = time.perf_counter()
        ob.direct_jk(dens, tau, schwarz=q, bounds=bounds, bra_range=(b0, b0 + m), threads=threads)
        dt = time.perf_counter() - t0
        nq = ob.last_stats[0]
        log("cpu sample %d bra pairs: %.2f s" % (m, dt))
        if dt >= seconds or b0 + m >= npairs:
            break
        grow = max(2.0, min(8.0, seconds / max(dt, 1e-3)))
        m = min(npairs - b0, int(m * grow) + 1)
    return {"value": nq / dt, "unit": "quartets/s", "cores": threads, "kind": "port",
            "sample": "canonical bra pairs [%d, %d) of %d (%d quartets, %.1f s), both densities, "
                      "J and K, tau=1e-12" % (b0, b0 + m, npairs, nq, dt)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="use the built-in class-to-kernel schedule instead of tuning it")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device once and exit: the reference "
                         "kernel of known traffic that scales the profiler's FETCH_SIZE and "
                         "WRITE_SIZE readings")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(the J/K all-reduce design is documented, not built); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    import torch
    if args.copy_probe:
        n = args.copy_probe * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log("copied %d MiB" % args.copy_probe)
        return
    import integral_amd as ig
    from integral_amd import synthetic

    shells, dens, tau = synthetic.direct_jk_config()
    basis = synthetic.to_basis(shells)
    nao = basis.nao()
    t0 = time.perf_counter()
    dc = ig.DirectContractor(basis)
    plan_s = time.perf_counter() - t0
    dev = torch.device("cuda:0")
    D = torch.tensor(np.stack(dens), dtype=torch.float64, device=dev).contiguous()
    J = torch.zeros_like(D)
    K = torch.zeros_like(D)
    stream = torch.cuda.current_stream(dev)

    def step():
        J.zero_()
        K.zero_()
        dc.jk_device(D, J, K, threshold=tau, stream=stream)

    log("plan built in %.2f s (nao=%d)" % (plan_s, nao))
    tune_s = 0.0
    if not args.no_autotune:
        t0 = time.perf_counter()
        dc.autotune(D, threshold=tau, reps=2, stream=stream)
        tune_s = time.perf_counter() - t0
        log("autotuned class-to-kernel schedule in %.2f s" % tune_s)
    sched = {e["cls"]: e["variant"] for e in dc.schedule()}
    log("warmup")
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize(dev)
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / args.steps
    quartets, prim_quartets = dc.last_stats()
    log("timed %d steps: %.3f ms/step, %d quartets" % (args.steps, ms, quartets))

    # per-class attribution with HIP events on the launch stream
    dc.set_profiling(True)
    acc = {}
    nprof = max(2, min(args.steps, 5))
    for _ in range(nprof):
        step()
        torch.cuda.synchronize(dev)
        for r in dc.profile():
            a = acc.setdefault(r["cls"], {"cls": r["cls"], "quartets": r["quartets"],
                                          "prim_quartets": r["prim_quartets"], "ms": 0.0})
            a["ms"] += r["ms"] / nprof
    dc.set_profiling(False)
    recs = sorted(acc.values(), key=lambda r: -r["ms"])
    dom = recs[0]
    dom_flops = class_flops(dom)
    total_flops = sum(class_flops(r) for r in recs)
    kern_ms = sum(r["ms"] for r in recs)
    achieved = dom_flops / (dom["ms"] * 1e-3) / 1e12
    roofline = {
        "bound": "mfma", "pipe": "FP64 vector ALU (78.6 TFLOP/s, same as dense FP64 matrix)",
        "kernel": kernel_name(dom["cls"], sched.get(tuple(dom["cls"]), 0)),
        "achieved": round(achieved, 4), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
        "frac": round(achieved / FP64_PEAK_TFLOPS, 5), "traffic": None,
        "launch_ms": round(dom["ms"], 4), "launch_flops": dom_flops,
        "all_classes_achieved": round(total_flops / (kern_ms * 1e-3) / 1e12, 4),
        "all_classes_frac": round(total_flops / (kern_ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS, 5),
    }
    cpu = None
    log("profiled classes; dominant %s %.3f ms" % (dom["cls"], dom["ms"]))
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(shells, dens, tau, args.cpu_seconds)
    out = {
        "metric": METRIC,
        "value": quartets / (ms * 1e-3),
        "unit": "quartets/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded molecule and densities, no checkpoint)",
        "config": {"workload": "integral-direct J/K, 40 atoms in a 14 bohr cube, spherical "
                               "s,s,p,p,d (3,3,3,3,1 primitives), nao=%d, 2 densities, "
                               "tau=1e-12, FP64" % nao,
                   "nao": nao, "nshell": len(shells), "ndens": 2, "tau": tau,
                   "quartets": quartets, "prim_quartets": prim_quartets,
                   "jk_build_s": ms * 1e-3, "plan_build_s": plan_s,
                   "schedule": "autotuned" if not args.no_autotune else "built-in",
                   "autotune_s": tune_s, "parallelism": "1 GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "classes": [{"cls": "".join("spdfghi"[l] for l in r["cls"]), "quartets": r["quartets"],
                     "variant": ["team", "register"][sched.get(tuple(r["cls"]), 0)],
                     "ms": round(r["ms"], 4)} for r in recs],
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
