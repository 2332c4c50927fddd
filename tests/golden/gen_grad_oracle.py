"""Writes tests/golden/grad_c4_oracle.json: the oracle's density-contracted
ERI gradient (the C restatement of direct_grad.rs:171-392, every canonical
quartet, all four centres differentiated) for the nao 390 energy-and-forces
workload (synthetic.gradient_config: 30 atoms, spherical s,s,p,p,d, one
density D), with the terms of the J/K energy E = 1/2 <D, J[D]> - 1/4 <D, K[D]>
(Coulomb (0.5, D, D), exchange (-0.25, D, D)) at tau 1e-12 -- the forces
bench.py's gradient workload times.  Runs on the host cores (ORACLE_THREADS)."""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from integral_amd import synthetic  # noqa: E402
from oracle import oracle as O  # noqa: E402

TAU = 1e-12
COULOMB, EXCHANGE = 0.5, -0.25


def main():
    shells, D = synthetic.gradient_config()
    ob = O.Basis(shells)
    t = time.time()
    F = ob.direct_grad([(COULOMB, D, D)], [(EXCHANGE, D, D)], TAU)
    dt = time.time() - t
    rec = {"workload": "gradient_config (nao 390, 30 atoms)", "tau": TAU,
           "coulomb": COULOMB, "exchange": EXCHANGE, "forces": F.tolist(),
           "quartets": ob.last_stats[0], "oracle_seconds": round(dt, 1),
           "threads": O.default_threads()}
    with open(os.path.join(HERE, "grad_c4_oracle.json"), "w") as f:
        json.dump(rec, f, indent=1)
    print("forces of %d atoms, %d quartets, %.1f s" % (F.shape[0], rec["quartets"], dt))


if __name__ == "__main__":
    main()
