"""The erf(omega r)/r operator on the ERI paths (Basis::eri_kernel,
EriBuilder::kernel, Basis::eri_2c_kernel / eri_3c_block_kernel): the oracle
has no attenuated integrals, so the values are pinned by
- the closed form for s functions: the Coulomb (ss|ss) expression with the
  Boys argument T -> s T and a factor sqrt(s), s = omega^2 / (rho + omega^2)
  (the attenuated-kernel transform of integrals.rs:304-321), after the same
  closed form with s = 1 has been checked against the Coulomb integrals;
- the limit omega -> infinity, where erf(omega r)/r -> 1/r, on s..f shells;
- the builder's bra-range fills against the dense tensor of the same call."""
import math

import numpy as np
import pytest

import integral_amd as ig
from integral_amd import synthetic
from tests import bases as B

pytestmark = pytest.mark.gpu


def note(msg):
    print("[case] " + msg, flush=True)


def boys0(x):
    return 1.0 - x / 3.0 + x * x / 10.0 if x < 1e-8 else 0.5 * math.sqrt(math.pi / x) * math.erf(math.sqrt(x))


def s_prim(a, A, b, B_, c, C, d, D, omega=None):
    """(ab|cd) of four unnormalised s primitives exp(-a|r-A|^2)..., with the
    operator 1/r (omega None) or erf(omega r)/r."""
    A, B_, C, D = (np.asarray(v, dtype=float) for v in (A, B_, C, D))
    p, q = a + b, c + d
    P, Q = (a * A + b * B_) / p, (c * C + d * D) / q
    kab = math.exp(-a * b / p * float(np.dot(A - B_, A - B_)))
    kcd = math.exp(-c * d / q * float(np.dot(C - D, C - D)))
    rho = p * q / (p + q)
    T = rho * float(np.dot(P - Q, P - Q))
    sc = 1.0 if omega is None else omega * omega / (rho + omega * omega)
    return 2.0 * math.pi ** 2.5 / (p * q * math.sqrt(p + q)) * kab * kcd * math.sqrt(sc) * boys0(sc * T)


def norm_s(a):
    return (2.0 * a / math.pi) ** 0.75


S_SHELLS = [(0.0, 0.0, 0.0, 1.3), (0.9, -0.4, 0.6, 0.45), (-0.7, 1.1, -0.3, 2.2),
            (0.2, 0.5, -0.9, 0.8)]


def s_basis():
    return [(0, (x, y, z), [a], [1.0], True) for x, y, z, a in S_SHELLS]


def s_tensor(omega):
    n = len(S_SHELLS)
    t = np.zeros((n, n, n, n))
    for i in range(n):
        for j in range(n):
            for k in range(n):
                for l in range(n):
                    si, sj, sk, sl = (S_SHELLS[m] for m in (i, j, k, l))
                    nn = norm_s(si[3]) * norm_s(sj[3]) * norm_s(sk[3]) * norm_s(sl[3])
                    t[i, j, k, l] = nn * s_prim(si[3], si[:3], sj[3], sj[:3], sk[3], sk[:3],
                                                sl[3], sl[:3], omega)
    return t.ravel()


def test_s_closed_form_coulomb_then_erf():
    basis = synthetic.to_basis(s_basis())
    note("s closed form, Coulomb")
    assert np.max(np.abs(basis.eri() - s_tensor(None))) < 1e-13
    for w in (0.3, 1.7):
        note("s closed form, erf omega=%g" % w)
        g = basis.eri_kernel(ig.EriKernel.erf(w))
        assert np.max(np.abs(g - s_tensor(w))) < 1e-13
        assert np.max(np.abs(g)) < np.max(np.abs(s_tensor(None)))


@pytest.mark.parametrize("sph", [False, True])
def test_erf_large_omega_is_coulomb(sph):
    """s..f shells (the f class pairs run on the runtime-l ERI kernel)."""
    shells = B.grad_basis(sph, max_l=3)
    basis = synthetic.to_basis(shells)
    note("erf omega=1e4 vs Coulomb sph=%s" % sph)
    c = basis.eri()
    e = basis.eri_kernel(ig.EriKernel.erf(1.0e4))
    assert np.max(np.abs(e - c)) <= 1e-7 * np.max(np.abs(c))
    e2 = basis.eri_kernel(ig.EriKernel.erf(0.5))
    assert np.max(np.abs(e2 - c)) > 1e-3 * np.max(np.abs(c))  # the operator did change


def test_erf_builder_fills_match_dense():
    shells = B.grad_basis(True, max_l=3)
    basis = synthetic.to_basis(shells)
    k = ig.EriKernel.erf(0.6)
    dense = basis.eri_kernel(k)
    eb = basis.eri_builder().kernel(k)
    out = np.zeros(eb.output_len())
    npair = len(eb.bra_pairs())
    note("erf builder fills in 3 ranges, %d bra pairs" % npair)
    for b0, b1 in ((0, npair // 3), (npair // 3, 2 * npair // 3), (2 * npair // 3, npair)):
        eb.fill_range(b0, b1, out)
    assert np.max(np.abs(out - dense)) < 1e-14
    assert np.array_equal(eb.build(), dense)
    with pytest.raises(ig.IntegralError):
        basis.eri_kernel(ig.EriKernel.erf(float("inf")))


def test_df_erf_closed_form_and_limit():
    aux = synthetic.to_basis(s_basis())
    n = len(S_SHELLS)
    # (P|Q): the unit-s dummy (exponent 0, coefficient 1) on each side
    ref = {}
    for w in (None, 0.7):
        m = np.zeros((n, n))
        for p in range(n):
            for q in range(n):
                sp, sq = S_SHELLS[p], S_SHELLS[q]
                m[p, q] = norm_s(sp[3]) * norm_s(sq[3]) * s_prim(
                    sp[3], sp[:3], 0.0, sp[:3], sq[3], sq[:3], 0.0, sq[:3], w)
        ref[w] = m.ravel()
    note("2c closed form")
    assert np.max(np.abs(aux.eri_2c() - ref[None])) < 1e-13
    assert np.max(np.abs(aux.eri_2c_kernel(ig.EriKernel.erf(0.7)) - ref[0.7])) < 1e-13
    # 3c: s main functions against the closed form, then s..d main/aux at large omega
    main = synthetic.to_basis(s_basis()[:3])
    note("3c closed form")
    for w in (None, 0.7):
        k = ig.EriKernel.coulomb() if w is None else ig.EriKernel.erf(w)
        blk = main.eri_3c_block_kernel(aux, 1, 0, 2, k)
        si, sj, sp = S_SHELLS[1], S_SHELLS[0], S_SHELLS[2]
        v = norm_s(si[3]) * norm_s(sj[3]) * norm_s(sp[3]) * s_prim(
            si[3], si[:3], sj[3], sj[:3], sp[3], sp[:3], 0.0, sp[:3], w)
        assert blk.shape == (1,) and abs(blk[0] - v) < 1e-13
    mb = synthetic.to_basis(B.grad_basis(True, max_l=2))
    ab = synthetic.to_basis(B.fd_basis())
    note("3c erf omega=1e4 vs Coulomb")
    c = mb.eri_3c_builder(ab).build()
    eb = mb.eri_3c_builder(ab)
    eb._kernel = ig.EriKernel.erf(1.0e4)
    e = eb.build()
    assert np.max(np.abs(e - c)) <= 1e-7 * np.max(np.abs(c))
    assert np.max(np.abs(mb.eri_2c_kernel(ig.EriKernel.erf(1.0e4)) - mb.eri_2c())) <= (
        1e-7 * np.max(np.abs(mb.eri_2c())))
