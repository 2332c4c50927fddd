"""Oracle J/K of the full-size direct J/K workloads, reduced to a committed
fixture the GPU tests check against at the workloads' real sizes:

- jk_c2_oracle.json: the nao 520 spherical s,s,p,p,d molecule (40 atoms), two
  densities, tau 1e-12 (integral_amd.synthetic.direct_jk_config);
- jk_c5_oracle.json: the nao 565 molecule with s..g shells and an l = 6
  shell on every fourth atom, one density, tau 1e-12 (high_l_config).

The oracle is the C restatement of direct.rs:131-218 (every canonical quartet
by the reference's own screening predicate).  For the high-l workload it runs
on the same shells listed by exponent (ascending; one primitive each), and
the matrices are permuted back: the reference's engines build each canonical
pair (i >= j) on shell i and transfer momentum to shell j, and when j is the
tight shell of a distant high-l pair that transfer cancels catastrophically
(12-28% of an (i i|i i) block, 0.3% of a (d i|d i) one, 5e-7 of an (i g|i g)
one, against 40-digit McMurchie-Davidson values); with the tighter shell as
shell i every block checked agrees with McMurchie-Davidson to about 5e-14.
The J/K are the same mathematical quantities in either order.  A fixture keeps, per matrix,
a deterministic sample of elements (strided, plus the largest) and 32 bilinear
probes u_k^T M v_k with seeded N(0,1) vectors -- a probe sums every element,
so a wrong element anywhere moves it.  Also the screened quartet count.
Run from the repository root: python tests/golden/gen_jk_oracle.py [c2|c5]
"""
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from integral_amd import synthetic  # noqa: E402
from oracle import oracle as O  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
NSAMPLE = 1500
NPROBE = 32


def probes(n, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((NPROBE, n)), rng.standard_normal((NPROBE, n))


def reduce(M, n, seed):
    M = np.asarray(M).reshape(n, n)
    flat = M.ravel()
    idx = np.unique(np.concatenate([np.arange(0, n * n, n * n // NSAMPLE + 1),
                                    [int(np.argmax(np.abs(flat)))]]))
    U, V = probes(n, seed)
    return {"index": [int(i) for i in idx], "value": [float(flat[i]) for i in idx],
            "probe": [float(u @ M @ v) for u, v in zip(U, V)], "max_abs": float(np.max(np.abs(M)))}


def main(which):
    if which == "c2":
        shells, dens, tau = synthetic.direct_jk_config()
        order = list(range(len(shells)))
    else:
        shells, D = synthetic.high_l_config()
        dens, tau = [D], 1e-12
        order = sorted(range(len(shells)), key=lambda k: (max(shells[k][2]), k))
    nf = [2 * s[0] + 1 if s[4] else (s[0] + 1) * (s[0] + 2) // 2 for s in shells]
    off = np.concatenate([[0], np.cumsum(nf)])
    perm = np.concatenate([np.arange(off[k], off[k + 1]) for k in order])  # new AO -> old AO
    ob = O.Basis([shells[k] for k in order])
    n = ob.nao
    t0 = time.time()
    J, K = ob.direct_jk([np.asarray(d)[np.ix_(perm, perm)] for d in dens], tau)
    inv = np.empty_like(perm)
    inv[perm] = np.arange(n)
    J = [np.asarray(M).reshape(n, n)[np.ix_(inv, inv)] for M in J]
    K = [np.asarray(M).reshape(n, n)[np.ix_(inv, inv)] for M in K]
    print("%s: nao %d, %d quartets, %.1f s" % (which, n, ob.last_stats[0], time.time() - t0),
          flush=True)
    out = {"workload": which, "nao": n, "tau": tau, "quartets": ob.last_stats[0],
           "probe_seed": 20261300, "J": [], "K": []}
    for k in range(len(dens)):
        out["J"].append(reduce(J[k], n, 20261300 + 2 * k))
        out["K"].append(reduce(K[k], n, 20261301 + 2 * k))
    with open(os.path.join(HERE, "jk_%s_oracle.json" % which), "w") as f:
        json.dump(out, f)


if __name__ == "__main__":
    for w in (sys.argv[1:] or ["c2", "c5"]):
        main(w)
