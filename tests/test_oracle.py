"""CPU suite: pins the oracle (the C restatement used as the parity checker)
against golden vectors and laws that do not depend on it.

Every check here would fail for a plausible mistake in the oracle: a dropped
recurrence term, a wrong sign, a transposed operand, a wrong normalisation.
"""
import json
import math
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests import bases as B

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)


# ---- Boys (math/boys.rs) ----------------------------------------------------

def test_boys_mpmath_golden():
    worst = 0.0
    for m, t, v in load("boys_mpmath.json")["rows"]:
        f = O.boys(m, t)[m]
        if v > 1e-300:
            worst = max(worst, abs(f - v) / v)
    assert worst < 1e-14, worst


def test_boys_known_values():
    # F_m(0) = 1/(2m+1), F_0(1) (boys.rs:467-520)
    for m in range(25):
        assert O.boys(m, 0.0)[m] == pytest.approx(1.0 / (2 * m + 1), rel=1e-15)
    assert abs(O.boys(0, 1.0)[0] - 0.7468241328124271) < 1e-15


def test_boys_large_t_closed_form_and_recurrence():
    for t in (95.0, 150.0, 1e3):
        f = O.boys(12, t)
        for m in range(13):
            closed = math.prod(range(1, 2 * m, 2)) / 2 ** (m + 1) * math.sqrt(math.pi) * t ** -(m + 0.5)
            assert abs(f[m] - closed) / closed < 1e-13
    for t in (0.3, 4.0, 21.0, 60.0):  # upward recurrence F_{m+1} = ((2m+1) F_m - e^-T) / 2T
        f = O.boys(10, t)
        for m in range(6):
            up = ((2 * m + 1) * f[m] - math.exp(-t)) / (2 * t)
            assert abs(up - f[m + 1]) / f[m + 1] < 1e-10


def test_boys_table_matches_reference_path():
    for m in (0, 3, 9, 17, 24):
        for t in np.linspace(0.0, 92.9, 37):
            a, r = O.boys(m, t), O.boys(m, t, reference=True)
            assert np.max(np.abs(a - r) / np.abs(r)) < 1e-13


# ---- Rys (math/rys.rs) --------------------------------------------------------

def test_rys_mpmath_golden():
    worst = 0.0
    for row in load("rys_mpmath.json")["rows"]:
        for ref in (False, True):
            x, w = O.rys(row["n"], row["T"], reference=ref)
            worst = max(worst, np.max(np.abs(x - row["roots"])), np.max(np.abs(w - row["weights"])))
    assert worst < 5e-14, worst


def test_rys_table_cell_counts_match_reference():
    # rys_tables.rs:12-25 publishes n_sub per n for the same recipe
    expected = [9, 8, 8, 8, 9, 10, 10, 11, 12, 13, 13, 14, 15]
    for n in range(1, 14):
        nsub, worst = O.rys_table_info(n)
        assert nsub == expected[n - 1]
        assert worst <= 2e-14


@pytest.mark.parametrize("n", [1, 2, 5, 9, 13])
def test_rys_moments_and_invariants(n):
    for t in (0.0, 0.7, 11.0, 33.0 + 2.5 * (2 * n - 1) - 0.5, 140.0):
        x, w = O.rys(n, t)
        assert np.all(x > 0) and np.all(x < 1) and np.all(np.diff(x) > 0) and np.all(w > 0)
        F = O.boys(2 * n - 1, t)
        for k in range(2 * n):
            # high moments are tiny: relative error grows, absolute stays ~1e-16
            assert abs(np.sum(w * x ** k) - F[k]) <= 1e-10 * F[k] + 1e-14 * F[0]
    x, w = O.rys(1, 2.0)
    F = O.boys(1, 2.0)
    assert abs(x[0] - F[1] / F[0]) < 1e-14 and abs(w[0] - F[0]) < 1e-14


# ---- c2s (math/solid_harmonics.rs) -------------------------------------------

def test_c2s_matches_reference_tables():
    g = load("c2s_reference.json")
    signed = iter(g["signed"])
    fixed = iter(g["sign_fixed"])
    for l in range(7):
        M = O.c2s_matrix(l)
        for row in M:
            assert np.max(np.abs(row - np.array(next(signed)))) < 1e-12
            r = row.copy()
            nz = np.flatnonzero(np.abs(r) > 1e-12)
            if r[nz[0]] < 0:
                r = -r
            assert np.max(np.abs(r - np.array(next(fixed)))) < 1e-12


# ---- ERI engines (engine/rys.rs, engine/os_eri.rs) ---------------------------

def test_eri_blocks_match_mcmurchie_davidson():
    g = load("md_eri.json")
    shells = [(s[0], tuple(s[1]), s[2], s[3], False) for s in g["shells"]]
    ob = O.Basis(shells)
    for q in g["quartets"]:
        ref = np.array(q["block"])
        for eng in (O.ENGINE_OS, O.ENGINE_RYS, O.ENGINE_AUTO):
            v = ob.eri_block(*q["shells"], engine=eng, cartesian=True).ravel()
            assert np.max(np.abs(v - ref)) < 1e-13 * max(1.0, np.max(np.abs(ref))), (q, eng)


def test_ssss_closed_form():
    """(ss|ss) = 2 pi^{5/2}/(p q sqrt(p+q)) K_ab K_cd F_0(T), F_0 from erf
    (engine/rys.rs:345-373)."""
    ex = [0.7, 1.3, 0.45, 2.2]
    cen = [np.zeros(3), np.array([0.3, 0.2, -0.4]), np.array([1.1, -0.6, 0.5]),
           np.array([0.9, 0.4, 1.0])]
    ob = O.Basis([(0, tuple(c), [e], [1.0], False) for e, c in zip(ex, cen)])
    N = lambda x: (2 * x / math.pi) ** 0.75
    a, b, c, d = ex
    p, q = a + b, c + d
    P = (a * cen[0] + b * cen[1]) / p
    Q = (c * cen[2] + d * cen[3]) / q
    kab = math.exp(-a * b / p * np.sum((cen[0] - cen[1]) ** 2))
    kcd = math.exp(-c * d / q * np.sum((cen[2] - cen[3]) ** 2))
    T = p * q / (p + q) * np.sum((P - Q) ** 2)
    f0 = 0.5 * math.sqrt(math.pi / T) * math.erf(math.sqrt(T))
    ref = 2 * math.pi ** 2.5 / (p * q * math.sqrt(p + q)) * kab * kcd * f0 * N(a) * N(b) * N(c) * N(d)
    for eng in (O.ENGINE_OS, O.ENGINE_RYS):
        g = ob.eri_block(0, 1, 2, 3, engine=eng).ravel()[0]
        assert abs(g - ref) / ref < 1e-14


def test_dense_eri_eightfold_symmetry_and_engines():
    shells = B.grad_basis(True)
    ob = O.Basis(shells)
    n = ob.nao
    e = ob.eri(O.ENGINE_AUTO).reshape(n, n, n, n)
    for perm in [(1, 0, 2, 3), (0, 1, 3, 2), (2, 3, 0, 1), (3, 2, 1, 0)]:
        assert np.max(np.abs(e - e.transpose(perm))) < 1e-14
    r = ob.eri(O.ENGINE_RYS).reshape(n, n, n, n)
    assert np.max(np.abs(e - r)) < 1e-11  # cross_engine.rs:45-46


def test_schwarz_bounds_every_element():
    shells = B.fd_basis()
    ob = O.Basis(shells)
    n = ob.nao
    e = np.abs(ob.eri().reshape(n, n, n, n))
    q = ob.schwarz().reshape(ob.nsh, ob.nsh)
    offs = np.cumsum([0] + [ob.nfunc(s) for s in range(ob.nsh)])
    for i in range(ob.nsh):
        for j in range(ob.nsh):
            for k in range(ob.nsh):
                for l in range(ob.nsh):
                    blk = e[offs[i]:offs[i + 1], offs[j]:offs[j + 1], offs[k]:offs[k + 1],
                            offs[l]:offs[l + 1]]
                    assert blk.max() <= q[i, j] * q[k, l] * (1 + 1e-12) + 1e-300


def test_eri_screened_skips_only_below_tau():
    shells = B.grad_basis(True) + [(1, (14.0, 0.0, 0.0), [0.9], [1.0], True)]
    ob = O.Basis(shells)
    full = ob.eri()
    scr, (total, skipped) = ob.eri_screened(1e-8)
    npair = ob.nsh * (ob.nsh + 1) // 2
    assert total == npair * (npair + 1) // 2 and skipped > 0
    assert np.max(np.abs(full - scr)) < 1e-8


# ---- direct J/K (direct.rs) ------------------------------------------------

def dense_jk(e, d):
    return np.einsum("abcd,cd->ab", e, d), np.einsum("abcd,bd->ac", e, d)


@pytest.mark.parametrize("sph", [False, True])
@pytest.mark.parametrize("engine", [O.ENGINE_OS, O.ENGINE_RYS])
def test_direct_jk_matches_dense(sph, engine):
    """direct_contractor.rs:30-91: non-symmetric densities, threshold 0."""
    shells = B.direct_contractor_basis(sph)
    ob = O.Basis(shells)
    n = ob.nao
    dens = B.direct_contractor_densities(n)
    e = ob.eri().reshape(n, n, n, n)
    J, K = ob.direct_jk(dens, 0.0, engine=engine)
    for k in range(2):
        Jr, Kr = dense_jk(e, dens[k])
        assert np.max(np.abs(J[k].reshape(n, n) - Jr)) < 1e-13
        assert np.max(np.abs(K[k].reshape(n, n) - Kr)) < 1e-13


def test_direct_jk_linear_symmetric_and_screened():
    shells = B.grad_basis(True) + B.shared_centers_basis()
    ob = O.Basis(shells)
    n = ob.nao
    d1, d2 = B.symmetric(n, 3), B.symmetric(n, 8)
    J, K = ob.direct_jk([d1, d2, 2.0 * d1 - d2], 0.0)
    assert np.max(np.abs(J[2] - (2 * J[0] - J[1]))) < 1e-12
    assert np.max(np.abs(K[2] - (2 * K[0] - K[1]))) < 1e-12
    Jm, Km = J[0].reshape(n, n), K[0].reshape(n, n)
    assert np.max(np.abs(Jm - Jm.T)) < 1e-13 and np.max(np.abs(Km - Km.T)) < 1e-13
    Js, Ks = ob.direct_jk([d1], 1e-6)
    assert 0 < np.max(np.abs(Js[0] - J[0])) < 1e-3


def test_direct_jk_bra_ranges_are_additive():
    shells = B.grad_basis(False)
    ob = O.Basis(shells)
    d = [B.symmetric(ob.nao, 2)]
    J, K = ob.direct_jk(d, 1e-12)
    np_ = ob.nsh * (ob.nsh + 1) // 2
    J1, K1 = ob.direct_jk(d, 1e-12, bra_range=(0, 7))
    J2, K2 = ob.direct_jk(d, 1e-12, bra_range=(7, np_))
    assert np.max(np.abs(J1[0] + J2[0] - J[0])) < 1e-13
    assert np.max(np.abs(K1[0] + K2[0] - K[0])) < 1e-13


# ---- gradient (direct_grad.rs, grad.rs) -----------------------------------

@pytest.mark.parametrize("sph", [False, True])
def test_direct_grad_matches_dense_reference(sph):
    shells = B.grad_basis(sph)
    ob = O.Basis(shells)
    n = ob.nao
    coul, exch = B.grad_terms(n)
    g = ob.direct_grad(coul, exch, 0.0)
    dense = ob.eri_grad_contract(B.dense_gamma(n, coul, exch))
    assert np.max(np.abs(g - dense)) < 1e-10  # direct_grad_contractor.rs:27
    assert np.max(np.abs(g.sum(axis=0))) < 1e-12
    assert np.max(np.abs(g)) > 1e-3


def shifted(shells, atom, axis, h):
    out = []
    for (l, c, e, k, s) in shells:
        if c == atom:
            c = tuple(c[i] + (h if i == axis else 0.0) for i in range(3))
        out.append((l, c, e, k, s))
    return out


def test_direct_grad_matches_finite_difference():
    """direct_grad_contractor.rs:304-330: h = 2e-4, tolerance 1e-6."""
    shells = B.fd_basis()
    ob = O.Basis(shells)
    n = ob.nao
    coul, exch = B.grad_terms(n)
    gamma = B.dense_gamma(n, coul, exch).ravel()
    g = ob.direct_grad(coul, exch, 0.0)
    atoms = []
    for s in shells:
        if s[1] not in atoms:
            atoms.append(s[1])
    h = 2e-4
    worst = 0.0
    for a, atom in enumerate(atoms):
        for ax in range(3):
            ep = np.dot(O.Basis(shifted(shells, atom, ax, h)).eri(), gamma)
            em = np.dot(O.Basis(shifted(shells, atom, ax, -h)).eri(), gamma)
            worst = max(worst, abs(g[a, ax] - (ep - em) / (2 * h)))
    assert worst < 1e-6, worst


def test_direct_grad_invariants():
    at = (0.3, -1.2, 0.8)
    one = [(0, at, [2.0, 0.5], [0.4, 0.7], False), (1, at, [0.9], [1.0], False),
           (2, at, [0.7], [1.0], True)]
    ob = O.Basis(one)
    coul, exch = B.grad_terms(ob.nao)
    for tau in (0.0, 1e-10):
        assert np.all(ob.direct_grad(coul, exch, tau) == 0.0)
    shells = B.grad_basis(True)
    ob = O.Basis(shells)
    coul, exch = B.grad_terms(ob.nao)
    both = ob.direct_grad(coul, exch, 0.0)
    split = ob.direct_grad(coul, [], 0.0) + ob.direct_grad([], exch, 0.0)
    assert np.max(np.abs(both - split)) < 1e-12
    np_ = ob.nsh * (ob.nsh + 1) // 2
    parts = ob.direct_grad(coul, exch, 0.0, bra_range=(0, 9)) + \
        ob.direct_grad(coul, exch, 0.0, bra_range=(9, np_))
    assert np.max(np.abs(parts - both)) < 1e-12


def test_direct_grad_screening_error_bounded():
    """direct_grad_contractor.rs:650-672: a finite threshold stays within 100 tau."""
    shells = B.grad_basis(True) + [(1, (14.0, 0.0, 0.0), [0.9], [1.0], True),
                                   (0, (14.0, 0.0, 0.0), [1.4, 0.4], [0.5, 0.6], True)]
    ob = O.Basis(shells)
    p = B.symmetric(ob.nao, 5)
    t = [(1.0, p, p)]
    exact = ob.direct_grad(t, t, 0.0)
    scr = ob.direct_grad(t, t, 1e-12)
    assert np.max(np.abs(exact - scr)) < 100 * 1e-12


def test_full_size_gradient_fixture_is_consistent():
    """tests/golden/grad_c4_oracle.json (the oracle's forces for the nao 390
    energy-and-forces workload): 30 atoms of 3 components, translation
    invariant (forces sum to zero, as test_direct_grad_invariants shows for
    the oracle at small size), and made for the workload's energy terms."""
    import json
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                        "grad_c4_oracle.json")
    with open(path) as f:
        rec = json.load(f)
    F = np.array(rec["forces"])
    assert F.shape == (30, 3)
    assert (rec["coulomb"], rec["exchange"], rec["tau"]) == (0.5, -0.25, 1e-12)
    assert np.max(np.abs(F.sum(axis=0))) < 1e-11 * max(1.0, np.max(np.abs(F)))
    assert np.max(np.abs(F)) > 0.1
