"""The oracle's restatement of PackedEriBuilder (packed_eri.rs) pinned by the
reference's own packed-ERI test assertions (crates/integral/tests/
packed_eri.rs): the AO-pair list, agreement with the dense tensor through it,
and the screening contract."""
import numpy as np
import pytest

from oracle import oracle as O


def ref_basis(spherical):
    """The basis of packed_eri.rs:3-11 (tests)."""
    return [
        (0, (0.0, 0.0, 0.0), [1.2, 0.35], [0.7, 0.4], spherical),
        (1, (0.2, -0.1, 0.5), [0.8], [1.0], spherical),
        (1, (0.2, -0.1, 0.5), [0.3], [1.0], spherical),
        (2, (-0.3, 0.4, -0.2), [0.6], [1.0], spherical),
        (3, (0.1, 0.3, 0.7), [0.5], [1.0], spherical),
    ]


@pytest.mark.parametrize("sph", [False, True])
def test_oracle_packed_layout_and_values(sph):
    ob = O.Basis(ref_basis(sph))
    n = ob.nao
    pairs, start = ob.packed_ao_pairs()
    assert len(pairs) == n * (n + 1) // 2 == start[-1]
    seen = set()
    for p, q in pairs:  # packed_eri.rs (tests):20-25
        assert p >= q and (p, q) not in seen
        seen.add((p, q))
    dense = ob.eri()
    packed = ob.packed_eri(0.0)
    assert packed.size == len(pairs) * (len(pairs) + 1) // 2
    peak = np.max(np.abs(dense))
    worst = 0.0
    for t, (p, q) in enumerate(pairs):
        for u, (r, s) in enumerate(pairs[:t + 1]):
            worst = max(worst, abs(packed[t * (t + 1) // 2 + u] - dense[((p * n + q) * n + r) * n + s]))
    assert worst < 1e-11 * max(peak, 1.0)
    # shell-pair blocking: every canonical pair's AO pairs are contiguous and
    # in (a, b) row-major order
    nf = [ob.nfunc(s) for s in range(ob.nsh)]
    k = 0
    for i in range(ob.nsh):
        for j in range(i + 1):
            assert start[k + 1] - start[k] == (nf[i] * (nf[i] + 1) // 2 if i == j else nf[i] * nf[j])
            k += 1


def test_oracle_packed_screening_contract():
    """packed_eri.rs (tests):63-86: screening only drops integrals at or below
    the bound, and the far-apart shells make it drop some."""
    shells = [
        (0, (0.0, 0.0, 0.0), [1.1], [1.0], False),
        (1, (0.0, 0.0, 0.4), [0.7], [1.0], False),
        (0, (0.0, 0.0, 5.0), [1.3], [1.0], False),
        (1, (0.0, 0.0, 5.4), [0.9], [1.0], False),
    ]
    ob = O.Basis(shells)
    exact = ob.packed_eri(0.0)
    tau = 1e-9
    screened = ob.packed_eri(tau)
    diff = exact != screened
    assert np.all(screened[diff] == 0.0)
    assert np.all(np.abs(exact[diff]) <= tau)
    assert np.count_nonzero(diff) > 0
