"""GPU density fitting (ig_eri_2c, ig_schwarz_aux_bounds, ig_eri_3c_fill)
against the oracle's restatement of df.rs and the reference's own DF test
assertions (crates/integral/tests/df.rs).  Bar: 1e-12 absolute."""
import math

import numpy as np
import pytest

import integral_amd as ig
from integral_amd import synthetic
from oracle import oracle as O
from tests.test_df_oracle import aux_basis, main_basis, pref, s_norm

pytestmark = pytest.mark.gpu

TOL = 1e-12


def note(msg):
    print("[case] " + msg, flush=True)


def test_df_closed_forms():
    note("df closed forms")
    a, b = 0.9, 1.7
    m = synthetic.to_basis([(0, (0.0, 0.0, 0.0), [a], [1.0], False),
                            (0, (0.0, 0.0, 0.0), [b], [1.0], False)]).eri_2c()
    assert abs(m[1] - s_norm(a) * s_norm(b) * pref(a, b)) < 1e-12 * max(1.0, abs(m[1]))
    assert abs(m[0] - s_norm(a) ** 2 * pref(a, a)) < 1e-12 * max(1.0, abs(m[0]))
    assert m[1] == m[2]
    a, b, c = 0.8, 1.1, 0.6
    main = synthetic.to_basis([(0, (0.0, 0.0, 0.0), [a], [1.0], False),
                               (0, (0.0, 0.0, 0.0), [b], [1.0], False)])
    aux = synthetic.to_basis([(0, (0.0, 0.0, 0.0), [c], [1.0], False)])
    blk = main.eri_3c_block(aux, 0, 1, 0)
    p = a + b
    want = s_norm(a) * s_norm(b) * s_norm(c) * 2.0 * math.pi ** 2.5 / (p * c * math.sqrt(p + c))
    assert blk.size == 1 and abs(blk[0] - want) < 1e-12 * max(1.0, want)


def test_df_matches_oracle_and_builder_contract():
    note("df 2c/3c vs oracle")
    main_s, aux_s = main_basis(), aux_basis()
    main, aux = synthetic.to_basis(main_s), synthetic.to_basis(aux_s)
    m = aux.eri_2c()
    om = O.df_eri_2c(aux_s)
    assert np.max(np.abs(m - om)) < TOL
    n = aux.nao()
    mm = m.reshape(n, n)
    assert np.array_equal(mm, mm.T)
    np.linalg.cholesky(mm)
    qp = aux.schwarz_aux_bounds()
    assert np.max(np.abs(qp - O.df_schwarz_aux(aux_s))) < TOL
    b = main.eri_3c_builder(aux)
    t = b.build()
    assert t.size == b.output_len() == main.nao() ** 2 * n
    assert np.max(np.abs(t - O.df_eri_3c(main_s, aux_s))) < TOL
    # tasks in reverse order reproduce the build bit for bit
    out = np.zeros(b.output_len())
    tasks = b.partition(out)
    assert [tk.bra() for tk in tasks] == b.bra_pairs()
    for tk in reversed(tasks):
        b.fill(tk)
    assert np.array_equal(out, t)
    # one block through the per-shell entry
    nao = main.nao()
    blk = main.eri_3c_block(aux, 3, 1, 2)
    t3 = t.reshape(nao, nao, n)
    offs = np.concatenate([[0], np.cumsum([s.n_func() for s in main.shells()])])
    aoffs = np.concatenate([[0], np.cumsum([s.n_func() for s in aux.shells()])])
    assert np.array_equal(blk, t3[offs[3]:offs[4], offs[1]:offs[2], aoffs[2]:aoffs[3]].ravel())
    with pytest.raises(ig.UsageError):
        b.partition(np.zeros(7))


def test_df_spdf_bases():
    """An s,p,d,f main basis with a g aux shell: the runtime-l kernels."""
    main_s, _ = synthetic.molecule(20260121, 2, 3.0, [(0, 2), (1, 1), (2, 1), (3, 1)],
                                   (0.2, 20.0), True)
    aux_s = [(0, (0.0, 0.0, 0.0), [2.1], [1.0], True), (2, (0.5, 0.3, -0.2), [0.9], [1.0], True),
             (4, (-0.3, 0.2, 0.4), [1.2], [1.0], True), (1, (0.1, -0.4, 0.6), [0.7], [1.0], False)]
    note("df spdf main, g aux")
    main, aux = synthetic.to_basis(main_s), synthetic.to_basis(aux_s)
    t = main.eri_3c_builder(aux).build()
    want = O.df_eri_3c(main_s, aux_s)
    assert np.max(np.abs(t - want)) <= TOL * max(1.0, float(np.max(np.abs(want))))
    m = aux.eri_2c()
    assert np.max(np.abs(m - O.df_eri_2c(aux_s))) <= TOL * max(1.0, float(np.max(np.abs(m))))
