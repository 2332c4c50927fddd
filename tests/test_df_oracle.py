"""The oracle's restatement of density fitting (df.rs: 2- and 3-center
integrals with a unit-s dummy) pinned by the reference's own DF test
assertions (crates/integral/tests/df.rs): closed forms for concentric s
functions, the tiny-exponent 4-center limit, exact symmetry, positive
definiteness and the aux Schwarz bound."""
import math

import numpy as np

from oracle import oracle as O


def s_norm(alpha):
    return (2.0 * alpha / math.pi) ** 0.75


def main_basis():
    """df.rs (tests):32-40."""
    return [
        (0, (0.0, 0.0, 0.0), [1.4, 0.5], [0.6, 0.5], False),
        (1, (0.6, -0.3, 0.2), [0.9], [1.0], False),
        (1, (0.6, -0.3, 0.2), [0.4], [1.0], False),
        (2, (-0.4, 0.7, -0.1), [1.1], [1.0], True),
    ]


def aux_basis():
    """df.rs (tests):42-50."""
    return [
        (0, (0.0, 0.0, 0.0), [2.1], [1.0], False),
        (1, (0.6, -0.3, 0.2), [1.3], [1.0], False),
        (2, (-0.4, 0.7, -0.1), [0.8], [1.0], True),
        (0, (0.3, 0.1, -0.2), [1.9, 0.7], [0.4, 0.8], False),
    ]


def pref(x, y):
    return 2.0 * math.pi ** 2.5 / (x * y * math.sqrt(x + y))


def test_oracle_2c_and_3c_closed_forms():
    """df.rs (tests):53-92."""
    a, b = 0.9, 1.7
    m = O.df_eri_2c([(0, (0.0, 0.0, 0.0), [a], [1.0], False),
                     (0, (0.0, 0.0, 0.0), [b], [1.0], False)])
    assert abs(m[1] - s_norm(a) * s_norm(b) * pref(a, b)) < 1e-12 * max(1.0, abs(m[1]))
    assert abs(m[0] - s_norm(a) ** 2 * pref(a, a)) < 1e-12 * max(1.0, abs(m[0]))
    assert abs(m[3] - s_norm(b) ** 2 * pref(b, b)) < 1e-12 * max(1.0, abs(m[3]))
    assert m[1] == m[2]
    a, b, c = 0.8, 1.1, 0.6
    t = O.df_eri_3c([(0, (0.0, 0.0, 0.0), [a], [1.0], False),
                     (0, (0.0, 0.0, 0.0), [b], [1.0], False)],
                    [(0, (0.0, 0.0, 0.0), [c], [1.0], False)])
    p = a + b
    want = s_norm(a) * s_norm(b) * s_norm(c) * 2.0 * math.pi ** 2.5 / (p * c * math.sqrt(p + c))
    assert abs(t.reshape(2, 2, 1)[1, 0, 0] - want) < 1e-12 * max(1.0, want)


def test_oracle_3c_tiny_exponent_limit():
    """df.rs (tests):94-124: (ij|P) equals the 4-center (ij|P s_eps) with an
    eps = 1e-12 s shell of effective coefficient 1."""
    main, aux = main_basis(), aux_basis()
    eps = 1e-12
    t = O.df_eri_3c(main, aux)
    nm = len(main)
    ob = O.Basis(main)
    n = ob.nao
    naux = sum(2 * s[0] + 1 if s[4] else (s[0] + 1) * (s[0] + 2) // 2 for s in aux)
    t = t.reshape(n, n, naux)
    off = np.concatenate([[0], np.cumsum([ob.nfunc(s) for s in range(nm)])])
    aoff = 0
    for sp in aux:
        four = O.Basis(main + [sp, (0, sp[1], [eps], [1.0 / s_norm(eps)], False)])
        npf = four.nfunc(nm)
        for i in range(nm):
            for j in range(nm):
                want = four.eri_block(i, j, nm, nm + 1).reshape(ob.nfunc(i), ob.nfunc(j), npf)
                got = t[off[i]:off[i + 1], off[j]:off[j + 1], aoff:aoff + npf]
                peak = float(np.max(np.abs(want)))
                assert np.max(np.abs(got - want)) <= 1e-8 * peak + 1e-12
        aoff += npf


def test_oracle_2c_symmetric_positive_and_aux_bound():
    """df.rs (tests):126-222."""
    aux = aux_basis()
    m = O.df_eri_2c(aux)
    n = int(round(math.sqrt(m.size)))
    m = m.reshape(n, n)
    assert np.array_equal(m, m.T)
    assert np.all(np.diag(m) > 0.0)
    np.linalg.cholesky(m)  # positive definite
    qp = O.df_schwarz_aux(aux)
    main = main_basis()
    q = O.Basis(main).schwarz().reshape(len(main), len(main))
    t = O.df_eri_3c(main, aux)
    ob = O.Basis(main)
    nao = ob.nao
    t = t.reshape(nao, nao, n)
    off = np.concatenate([[0], np.cumsum([ob.nfunc(s) for s in range(len(main))])])
    aoff = np.concatenate([[0], np.cumsum([2 * s[0] + 1 if s[4] else (s[0] + 1) * (s[0] + 2) // 2
                                           for s in aux])])
    for i in range(len(main)):
        for j in range(len(main)):
            for p in range(len(aux)):
                blk = t[off[i]:off[i + 1], off[j]:off[j + 1], aoff[p]:aoff[p + 1]]
                assert np.max(np.abs(blk)) <= q[i, j] * qp[p] * (1 + 1e-10) + 1e-14
