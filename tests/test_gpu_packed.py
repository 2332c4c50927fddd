"""GPU PackedEriBuilder (ig_packed_eri_*) against the oracle's restatement of
packed_eri.rs and the reference's own test assertions
(crates/integral/tests/packed_eri.rs).  Bar: 1e-12 absolute."""
import numpy as np
import pytest

import integral_amd as ig
from integral_amd import synthetic
from oracle import oracle as O
from tests.test_packed_oracle import ref_basis

pytestmark = pytest.mark.gpu

TOL = 1e-12


def note(msg):
    print("[case] " + msg, flush=True)


@pytest.mark.parametrize("sph", [False, True])
def test_packed_matches_oracle_and_fills(sph):
    shells = ref_basis(sph)
    note("packed eri sph=%s" % sph)
    basis = synthetic.to_basis(shells)
    pb = basis.packed_eri_builder()
    ob = O.Basis(shells)
    opairs, _ = ob.packed_ao_pairs()
    assert pb.ao_pairs() == opairs
    assert pb.output_len() == len(opairs) * (len(opairs) + 1) // 2
    serial = pb.build(0.0)
    assert serial.size == pb.output_len()
    assert np.max(np.abs(serial - ob.packed_eri(0.0))) < TOL
    # against the GPU's own dense tensor through the AO pairs (the dense path
    # may evaluate a quartet with bra and ket exchanged: equal to round-off)
    n = basis.nao()
    dense = basis.eri()
    idx = [((p * n + q) * n + r) * n + s for t, (p, q) in enumerate(opairs)
           for (r, s) in opairs[:t + 1]]
    assert np.max(np.abs(serial - dense[idx])) < 1e-14
    # tasks filled in reverse order reproduce the serial build bit for bit
    out = np.zeros(pb.output_len())
    tasks = pb.partition(out)
    assert len(tasks) == len(pb.bra_pairs())
    for task in reversed(tasks):
        pb.fill(ig.PackedEriWorkspace(), task, 0.0)
    assert np.array_equal(out, serial)
    # a range of bra pairs at once
    b0, b1 = 3, 11
    off = sum(t.rows.size for t in tasks[:b0])
    ln = sum(t.rows.size for t in tasks[b0:b1])
    rows = np.zeros(ln)
    pb.fill_range(b0, b1, rows)
    assert np.array_equal(rows, serial[off:off + ln])


def test_packed_screening_and_underflow():
    shells = [
        (0, (0.0, 0.0, 0.0), [1.1], [1.0], False),
        (1, (0.0, 0.0, 0.4), [0.7], [1.0], False),
        (0, (0.0, 0.0, 5.0), [1.3], [1.0], False),
        (1, (0.0, 0.0, 5.4), [0.9], [1.0], False),
    ]
    note("packed eri screening")
    pb = synthetic.to_basis(shells).packed_eri_builder()
    exact = pb.build(0.0)
    tau = 1e-9
    screened = pb.build(tau)
    diff = exact != screened
    assert np.all(screened[diff] == 0.0) and np.all(np.abs(exact[diff]) <= tau)
    assert np.count_nonzero(diff) > 0
    ob = O.Basis(shells)
    assert np.max(np.abs(screened - ob.packed_eri(tau))) < TOL
    # two s shells 13 bohr apart: the mixed pair's bound underflows; the packed
    # and dense builds agree element for element, to round-off
    # (packed_eri.rs tests:88-115; the ~3e-38 integral there is below this
    # build's primitive-pair cut and is 0 in both)
    far = [(0, (0.0, 0.0, 0.0), [1.0], [1.0], False), (0, (0.0, 0.0, 13.0), [1.0], [1.0], False)]
    b = synthetic.to_basis(far)
    pk = b.packed_eri_builder()
    dense = b.eri()
    n = b.nao()
    pairs = pk.ao_pairs()
    packed = pk.build(0.0)
    for t, (p, q) in enumerate(pairs):
        for u, (r, s) in enumerate(pairs[:t + 1]):
            want = dense[((p * n + q) * n + r) * n + s]
            assert abs(packed[t * (t + 1) // 2 + u] - want) <= 1e-15 * abs(want)


def test_packed_spdf_molecule_and_bad_arguments():
    """A 3-atom s,p,d,f molecule plus a Cartesian f shell: the compiled and
    the runtime-l ERI kernels, spherical and Cartesian shells together."""
    shells, _ = synthetic.molecule(20260111, 3, 4.0, [(0, 2), (1, 2), (2, 1), (3, 1)],
                                   (0.1, 30.0), True)
    shells += [(3, (0.4, 1.1, -0.7), [0.8, 0.3], [0.6, 0.5], False)]
    note("packed eri spdf nsh=%d" % len(shells))
    pb = synthetic.to_basis(shells).packed_eri_builder()
    got = pb.build(0.0)
    want = O.Basis(shells).packed_eri(0.0)
    # the f-class bar of the other high-l tests: 1e-12 relative to the largest
    # integral (8.3 here; the oracle's own OS-vs-Rys spread on this basis is
    # 1.1e-12 absolute)
    assert np.max(np.abs(got - want)) <= TOL * max(1.0, float(np.max(np.abs(want))))
    with pytest.raises(ig.UsageError):
        pb.partition(np.zeros(3))
    with pytest.raises(ig.IntegralError):
        pb.fill_range(5, 2, np.zeros(1))


@pytest.mark.parametrize("sph", [False, True])
def test_eri_builder_fills_own_rows(sph):
    """EriBuilder (eri_builder.rs): tasks filled in reverse order assemble
    the eri() tensor to round-off, and one task writes only the rows of its
    bra pair (mu in i, nu in j, or the reverse)."""
    shells = ref_basis(sph)
    note("eri builder sph=%s" % sph)
    basis = synthetic.to_basis(shells)
    eb = basis.eri_builder()
    dense = basis.eri()
    assert np.max(np.abs(dense - O.Basis(shells).eri())) < TOL
    out = np.zeros(eb.output_len())
    tasks = eb.partition(out)
    assert len(tasks) == len(eb.bra_pairs())
    for task in reversed(tasks):
        eb.fill(task)
    assert np.max(np.abs(out - dense)) < 1e-14
    n = basis.nao()
    offs = np.concatenate([[0], np.cumsum([2 * s[0] + 1 if sph else (s[0] + 1) * (s[0] + 2) // 2
                                           for s in shells])])
    for k in (0, 4, len(tasks) - 1):
        i, j = eb.bra_pairs()[k]
        one = np.zeros(eb.output_len())
        eb.fill(eb.partition(one)[k])
        t = one.reshape(n, n, n, n)
        own = np.zeros((n, n), dtype=bool)
        own[offs[i]:offs[i + 1], offs[j]:offs[j + 1]] = True
        own[offs[j]:offs[j + 1], offs[i]:offs[i + 1]] = True
        assert np.all(t[~own] == 0.0)
        assert np.max(np.abs(t[own] - dense.reshape(n, n, n, n)[own])) < 1e-14
    with pytest.raises(ig.UsageError):
        eb.partition(np.zeros(5))
    with pytest.raises(ig.UsageError):  # omega must be finite and > 0
        eb.kernel(ig.EriKernel.erf(-0.4))
