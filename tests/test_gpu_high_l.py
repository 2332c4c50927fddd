"""GPU parity for shells above d: the runtime-l kernels (f through i, and the
gradient's raised tables up to l = 6) against the oracle.

The f-shell bases are the reference's own test inputs
(direct_contractor.rs:3-45, direct_grad_contractor.rs:32-120); the g/h/i
cases are small synthetic bases that put one high shell next to s and p.
Bar: 1e-12 relative to the largest entry (FP64 path)."""
import numpy as np
import pytest

import integral_amd as ig
from integral_amd import synthetic
from oracle import oracle as O
from tests import bases as B

pytestmark = pytest.mark.gpu


def note(msg):
    print("[case] " + msg, flush=True)


def close(g, o, rel=1e-12):
    scale = max(1.0, float(np.max(np.abs(o))))
    err = float(np.max(np.abs(np.asarray(g) - np.asarray(o))))
    assert err <= rel * scale, (err, scale)


def high_basis(l, sph):
    """One shell of angular momentum l between an s and a p shell."""
    return [
        (0, (0.0, 0.0, 0.0), [1.3, 0.4], [0.6, 0.5], sph),
        (l, (0.3, -0.2, 0.4), [0.7], [1.0], sph),
        (1, (-0.5, 0.6, -0.1), [0.9], [1.0], sph),
    ]


@pytest.mark.parametrize("sph", [False, True])
def test_schwarz_f_to_i(sph):
    shells = B.direct_contractor_basis(sph) + [
        (4, (0.4, 0.1, -0.3), [0.8], [1.0], sph),
        (5, (-0.2, 0.5, 0.6), [0.9], [1.0], sph),
        (6, (0.6, -0.4, 0.2), [1.1], [1.0], sph),
    ]
    note("schwarz f..i sph=%s" % sph)
    g = synthetic.to_basis(shells).schwarz_bounds()
    o = O.Basis(shells).schwarz()
    assert np.max(np.abs(g - o) / np.maximum(np.abs(o), 1e-300)) < 1e-12


@pytest.mark.parametrize("sph", [False, True])
def test_eri_dense_reference_basis_with_f(sph):
    shells = B.direct_contractor_basis(sph)
    note("eri dense s,p,d,f sph=%s" % sph)
    g = synthetic.to_basis(shells).eri()
    o = O.Basis(shells).eri()
    close(g, o)


@pytest.mark.parametrize("l", [4, 5, 6])
def test_eri_dense_g_h_i(l):
    shells = high_basis(l, True)
    note("eri dense l=%d spherical" % l)
    g = synthetic.to_basis(shells).eri()
    o = O.Basis(shells).eri()
    close(g, o)
    n = B.nfunc(shells)
    t = g.reshape(n, n, n, n)
    assert np.max(np.abs(t - t.transpose(2, 3, 0, 1))) <= 1e-14 * max(1.0, np.max(np.abs(t)))


@pytest.mark.parametrize("sph", [False, True])
def test_direct_jk_reference_basis_with_f(sph):
    """direct_contractor.rs:12-45: s, p, d, f with two non-symmetric densities."""
    shells = B.direct_contractor_basis(sph)
    n = B.nfunc(shells)
    dens = B.direct_contractor_densities(n)
    note("jk s,p,d,f sph=%s" % sph)
    dc = synthetic.to_basis(shells).direct_contractor()
    J, K = dc.jk(dens, 0.0)
    oJ, oK = O.Basis(shells).direct_jk(dens, 0.0)
    for x in range(2):
        close(J[x], oJ[x])
        close(K[x], oK[x])


@pytest.mark.parametrize("l", [4, 5, 6])
def test_direct_jk_g_h_i(l):
    shells = high_basis(l, l % 2 == 0)
    n = B.nfunc(shells)
    p = B.symmetric(n, l)
    note("jk l=%d" % l)
    dc = synthetic.to_basis(shells).direct_contractor()
    J, K = dc.jk([p], 1e-12)
    oJ, oK = O.Basis(shells).direct_jk([p], 1e-12)
    close(J[0], oJ[0])
    close(K[0], oK[0])


@pytest.mark.parametrize("sph", [False, True])
def test_grad_reference_basis_with_f(sph):
    """direct_grad_contractor.rs:32-120: f shell on the third centre."""
    shells = B.grad_basis(sph)
    n = B.nfunc(shells)
    coul, exch = B.grad_terms(n)
    note("grad s..f sph=%s" % sph)
    gc = synthetic.to_basis(shells).direct_grad_contractor()
    dens = gc.density([ig.GradDensityTerm(*t) for t in coul],
                      [ig.GradDensityTerm(*t) for t in exch])
    g = gc.contract(dens)
    o = O.Basis(shells).direct_grad(coul, exch, 0.0)
    close(g, o)
    assert np.max(np.abs(g.sum(axis=0))) < 1e-11


@pytest.mark.parametrize("l", [4, 5])
def test_grad_g_h(l):
    shells = high_basis(l, True)
    n = B.nfunc(shells)
    coul, exch = B.grad_terms(n)
    note("grad l=%d" % l)
    gc = synthetic.to_basis(shells).direct_grad_contractor()
    dens = gc.density([ig.GradDensityTerm(*t) for t in coul],
                      [ig.GradDensityTerm(*t) for t in exch])
    g = gc.contract(dens)
    o = O.Basis(shells).direct_grad(coul, exch, 0.0)
    close(g, o)
    assert np.max(np.abs(g.sum(axis=0))) < 1e-11
