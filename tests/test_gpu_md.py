"""GPU integrals against the independent McMurchie-Davidson vectors
(tests/golden/gen_md_ld.py, extended precision, checked against the 40-digit
mpmath set before writing):

- the erf(omega r)/r operator at finite omega (0.4 and 2.5) on s..h shells,
  through Basis::eri_kernel, elementwise at 1e-12 of max(1, peak): the pin the
  reference's own erf suite uses (crates/integral/tests/erf_kernel.rs:46-230);
- Coulomb blocks with g, h and i shells (l_total up to 16), through
  Basis::eri, at the same bar;
- the erf gradient of the dense two-particle-density contraction
  (grad.rs:493-540) against a Richardson-extrapolated central difference of
  the erf energy sum(Gamma * (mu nu|lam sig)) built from those same pinned
  integrals, at 1e-8 (the difference scheme's own error is ~1e-10 here).
"""
import json
import os

import numpy as np
import pytest

import integral_amd as ig
from integral_amd import synthetic
from tests import bases as B

pytestmark = pytest.mark.gpu

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAO_CAP = 50  # dense tensors of the quartet's own shells stay small (50^4 doubles)


def note(msg):
    print("[case] " + msg, flush=True)


def load(name):
    with open(os.path.join(HERE, name)) as f:
        return json.load(f)


def ncart(l):
    return (l + 1) * (l + 2) // 2


def quartet_block(shells, q, kernel):
    """Cartesian block (q0 q1|q2 q3) from the dense tensor of a basis made of
    the quartet's distinct shells only."""
    uniq = sorted(set(q))
    sub = [(shells[i][0], tuple(shells[i][1]), shells[i][2], shells[i][3], False) for i in uniq]
    off, n = {}, 0
    for i in uniq:
        off[i] = n
        n += ncart(shells[i][0])
    if n > NAO_CAP:
        return None
    basis = synthetic.to_basis(sub)
    t = (basis.eri() if kernel is None else basis.eri_kernel(kernel)).reshape(n, n, n, n)
    sl = [slice(off[i], off[i] + ncart(shells[i][0])) for i in q]
    return t[sl[0], sl[1], sl[2], sl[3]].ravel()


def test_erf_blocks_match_mcmurchie_davidson():
    g = load("md_erf.json")
    for rec in g["quartets"]:
        note("erf omega=%g quartet %s" % (rec["omega"], rec["shells"]))
        v = quartet_block(g["shells"], rec["shells"], ig.EriKernel.erf(rec["omega"]))
        assert v is not None and v.size == rec["n"]
        ref = np.array(rec["value"])
        err = np.max(np.abs(v[np.array(rec["index"])] - ref))
        assert err < 1e-12 * max(1.0, np.max(np.abs(ref))), (rec["shells"], rec["omega"], err)


def test_high_l_blocks_match_mcmurchie_davidson():
    g = load("md_eri_high.json")
    done = 0
    for rec in g["quartets"]:
        v = quartet_block(g["shells"], rec["shells"], None)
        if v is None:
            continue
        note("high-l quartet %s" % rec["shells"])
        ref = np.array(rec["value"])
        err = np.max(np.abs(v[np.array(rec["index"])] - ref))
        assert err < 1e-12 * max(1.0, np.max(np.abs(ref))), (rec["shells"], err)
        done += 1
    assert done >= 6


def _shifted(shells, atom_center, axis, h):
    out = []
    for l, c, e, k, sph in shells:
        c = list(c)
        if tuple(c) == atom_center:
            c[axis] += h
        out.append((l, tuple(c), e, k, sph))
    return out


def test_erf_gradient_matches_finite_difference_of_pinned_integrals():
    shells = B.grad_basis(True, max_l=2)
    n = B.nfunc(shells)
    gamma = np.random.default_rng(20261205).standard_normal((n, n, n, n)) * 0.1
    k = ig.EriKernel.erf(0.7)
    basis = synthetic.to_basis(shells)
    grad = basis.eri_grad_contract_kernel(gamma, k)
    atoms = []
    for s in shells:
        if tuple(s[1]) not in atoms:
            atoms.append(tuple(s[1]))
    assert grad.shape == (len(atoms), 3)

    def energy(sh):
        return float(np.dot(gamma.ravel(), synthetic.to_basis(sh).eri_kernel(k)))

    def central(a, x, h):
        return (energy(_shifted(shells, a, x, h)) - energy(_shifted(shells, a, x, -h))) / (2 * h)

    worst = 0.0
    for ia, a in enumerate(atoms):
        for x in range(3):
            note("erf gradient FD atom %d axis %d" % (ia, x))
            d1, d2 = central(a, x, 2e-3), central(a, x, 1e-3)
            fd = (4.0 * d2 - d1) / 3.0  # Richardson: O(h^4)
            worst = max(worst, abs(fd - grad[ia, x]))
    assert worst < 1e-8, worst
