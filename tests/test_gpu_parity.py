"""GPU parity: the HIP library (through its C ABI) against the oracle on the
same inputs.  Bar: 1e-12 absolute on ERIs, J/K and gradients (FP64 path)."""
import numpy as np
import pytest

import integral_amd as ig
from integral_amd import synthetic
from oracle import oracle as O
from tests import bases as B

pytestmark = pytest.mark.gpu

TOL = 1e-12


def gpu_basis(shells):
    return synthetic.to_basis(shells)


def note(msg):
    print("[case] " + msg, flush=True)


@pytest.mark.parametrize("sph", [False, True])
def test_schwarz_matches_oracle(sph):
    shells = B.direct_contractor_basis(sph, max_l=2) + B.fd_basis()
    note("schwarz sph=%s nsh=%d" % (sph, len(shells)))
    g = gpu_basis(shells).schwarz_bounds()
    o = O.Basis(shells).schwarz()
    assert np.max(np.abs(g - o) / np.maximum(np.abs(o), 1e-300)) < 1e-12


@pytest.mark.parametrize("sph", [False, True])
def test_eri_dense_small_matches_oracle(sph):
    shells = B.grad_basis(sph, max_l=2)
    note("eri dense sph=%s" % sph)
    g = gpu_basis(shells).eri()
    o = O.Basis(shells).eri()
    assert np.max(np.abs(g - o)) < TOL


def test_eri_dense_mixed_kinds():
    shells = B.fd_basis()
    note("eri dense mixed kinds")
    g = gpu_basis(shells).eri()
    o = O.Basis(shells).eri()
    assert np.max(np.abs(g - o)) < TOL


def test_eri_dense_config_c1():
    shells, _ = synthetic.dense_eri_config()
    note("eri dense 10-atom s,s,p Cartesian nao=50")
    gb = gpu_basis(shells)
    assert gb.nao() == 50
    g = gb.eri()
    o = O.Basis(shells).eri()
    assert np.max(np.abs(g - o)) < TOL
    n = 50
    t = g.reshape(n, n, n, n)
    assert np.max(np.abs(t - t.transpose(1, 0, 2, 3))) < 1e-14
    assert np.max(np.abs(t - t.transpose(2, 3, 0, 1))) < 1e-14


def test_eri_screened_stats_and_values():
    shells = B.grad_basis(True, max_l=2) + [(1, (14.0, 0.0, 0.0), [0.9], [1.0], True)]
    note("eri screened")
    gb, ob = gpu_basis(shells), O.Basis(shells)
    g, st = gb.eri_screened(1e-8)
    o, ost = ob.eri_screened(1e-8)
    assert (st["shell_quartets_total"], st["shell_quartets_skipped"]) == ost
    assert np.max(np.abs(g - o)) < TOL


@pytest.mark.parametrize("sph", [False, True])
def test_direct_jk_nonsymmetric_densities_tau0(sph):
    shells = B.direct_contractor_basis(sph, max_l=2)
    n = B.nfunc(shells)
    dens = B.direct_contractor_densities(n)
    note("jk nonsymmetric sph=%s" % sph)
    dc = gpu_basis(shells).direct_contractor()
    J, K = dc.jk(dens, 0.0)
    Jo, Ko = O.Basis(shells).direct_jk(dens, 0.0)
    for d in range(2):
        assert np.max(np.abs(J[d] - Jo[d])) < TOL
        assert np.max(np.abs(K[d] - Ko[d])) < TOL


def test_direct_jk_screened_symmetric():
    shells = (B.grad_basis(True, max_l=2) + B.shared_centers_basis()
              + [(1, (9.0, 0.0, 0.0), [0.9], [1.0], True),
                 (0, (9.0, 0.0, 0.0), [1.4, 0.4], [0.5, 0.6], True)])
    n = B.nfunc(shells)
    dens = [B.symmetric(n, 4), B.symmetric(n, 7) * 1e-3]
    note("jk screened symmetric n=%d" % n)
    dc = gpu_basis(shells).direct_contractor()
    ob = O.Basis(shells)
    for tau in (0.0, 1e-10, 1e-6):
        J, K = dc.jk(dens, tau)
        Jo, Ko = ob.direct_jk(dens, tau)
        assert dc.last_stats()[0] == ob.last_stats[0], tau
        for d in range(2):
            assert np.max(np.abs(J[d] - Jo[d])) < TOL
            assert np.max(np.abs(K[d] - Ko[d])) < TOL


def test_direct_jk_bra_pair_compat_and_ranges():
    shells = B.grad_basis(True, max_l=2)
    n = B.nfunc(shells)
    dens = [B.symmetric(n, 2)]
    note("jk per-bra compat")
    dc = gpu_basis(shells).direct_contractor()
    Jb, Kb = dc.jk(dens, 1e-12)
    J, K = np.zeros(n * n), np.zeros(n * n)
    bounds = dc.density_shell_bounds(dens)
    ws = ig.DirectWorkspace()
    for bra in reversed(dc.bra_pairs()):
        dc.accumulate_bra_pair(ws, bra, dens, bounds, 1e-12, True, [ig.DirectBuffers(J, K)])
    assert np.max(np.abs(J - Jb[0])) < TOL and np.max(np.abs(K - Kb[0])) < TOL
    np_ = len(dc.bra_pairs())
    J2, K2 = dc.jk(dens, 1e-12, bra_range=(0, np_ // 3))
    dc.jk(dens, 1e-12, bra_range=(np_ // 3, np_), J=J2, K=K2)
    assert np.max(np.abs(J2[0] - Jb[0])) < TOL and np.max(np.abs(K2[0] - Kb[0])) < TOL


@pytest.mark.parametrize("sph", [False, True])
def test_grad_matches_oracle(sph):
    shells = B.grad_basis(sph, max_l=2)
    n = B.nfunc(shells)
    coul, exch = B.grad_terms(n)
    note("grad sph=%s" % sph)
    gc = gpu_basis(shells).direct_grad_contractor()
    dens = gc.density([ig.GradDensityTerm(*t) for t in coul], [ig.GradDensityTerm(*t) for t in exch])
    g = gc.contract(dens)
    o = O.Basis(shells).direct_grad(coul, exch, 0.0)
    assert np.max(np.abs(g - o)) < TOL
    assert np.max(np.abs(g.sum(axis=0))) < 1e-12


def test_grad_mixed_and_shared_centres():
    for shells in (B.fd_basis(), B.shared_centers_basis()):
        n = B.nfunc(shells)
        coul, exch = B.grad_terms(n)
        note("grad basis nsh=%d" % len(shells))
        gc = gpu_basis(shells).direct_grad_contractor()
        dens = gc.density([ig.GradDensityTerm(*t) for t in coul],
                          [ig.GradDensityTerm(*t) for t in exch])
        g = gc.contract(dens)
        o = O.Basis(shells).direct_grad(coul, exch, 0.0)
        assert np.max(np.abs(g - o)) < TOL


def test_grad_bra_pairs_sum_in_any_order():
    shells = B.grad_basis(True, max_l=2)
    n = B.nfunc(shells)
    p = B.symmetric(n, 4)
    note("grad per-bra")
    gc = gpu_basis(shells).direct_grad_contractor()
    t = [ig.GradDensityTerm(1.0, p, p)]
    dens = gc.density(t, t)
    serial = gc.contract(dens, threshold=0.0)
    parts = [np.zeros((gc.natom(), 3)) for _ in range(2)]
    for k, bra in enumerate(reversed(gc.bra_pairs())):
        gc.accumulate_bra_pair(None, bra, dens, ig.EriKernel.coulomb(), 0.0, parts[k % 2])
    assert np.max(np.abs(parts[0] + parts[1] - serial)) < 1e-12


@pytest.mark.parametrize("variant", [0, 1, 2, 3, 4])
def test_direct_jk_every_kernel_variant(variant):
    """Every J/K kernel of every compiled class (team/LDS, lane/register,
    register with per-row LDS sums, the OS/HGP block for total l <= 2, and
    the team kernel with contracted-space transfers) against the oracle,
    forced through a loaded schedule (classes a variant cannot run take the
    register kernel)."""
    shells = (B.grad_basis(True, max_l=2) + B.fd_basis()
              + [(2, (1.3, -0.8, 0.9), [0.45], [1.0], True)])
    n = B.nfunc(shells)
    dens = B.direct_contractor_densities(n)
    note("jk variant %d nsh=%d" % (variant, len(shells)))
    dc = gpu_basis(shells).direct_contractor()
    sched = dc.schedule()
    dc.set_schedule([dict(e, variant=variant) for e in sched])
    assert all(e["variant"] == variant and e["tuned"] for e in dc.schedule())
    assert variant != 3 or any(sum(e["cls"]) <= 2 for e in sched)
    J, K = dc.jk(dens, 0.0)
    oJ, oK = O.Basis(shells).direct_jk(dens, 0.0)
    for x in range(2):
        assert np.max(np.abs(J[x] - oJ[x])) < TOL
        assert np.max(np.abs(K[x] - oK[x])) < TOL


@pytest.mark.parametrize("variant", [None, 0, 1, 2, 4])
def test_direct_jk_f_classes(variant):
    """The compiled f-shell J/K kernels (team kernels for f class pairs up to
    3600 elements, lane kernels up to 270, (ff|fd) and (ff|ff) on the team
    kernel with contracted-space transfers only) against the oracle at tau = 0
    on non-symmetric densities, spherical and Cartesian f shells."""
    shells, _ = synthetic.molecule(20260111, 3, 4.0, [(0, 2), (1, 2), (2, 1), (3, 1)],
                                   (0.1, 30.0), True)
    shells += [(3, (0.4, 1.1, -0.7), [0.8, 0.3], [0.6, 0.5], False),
               (1, (0.4, 1.1, -0.7), [0.5], [1.0], True)]
    n = B.nfunc(shells)
    dens = B.direct_contractor_densities(n)
    note("jk f classes variant %s nsh=%d nao=%d" % (variant, len(shells), n))
    dc = gpu_basis(shells).direct_contractor()
    sched = dc.schedule()
    assert any(max(e["cls"]) == 3 for e in sched)
    if variant is not None:
        dc.set_schedule([dict(e, variant=variant) for e in sched])
    J, K = dc.jk(dens, 0.0)
    oJ, oK = O.Basis(shells).direct_jk(dens, 0.0)
    for x in range(2):
        assert np.max(np.abs(J[x] - oJ[x])) < TOL
        assert np.max(np.abs(K[x] - oK[x])) < TOL


def test_autotune_keeps_results():
    import torch
    shells = B.grad_basis(True, max_l=2) + B.shared_centers_basis()
    n = B.nfunc(shells)
    p = B.symmetric(n, 9)
    note("autotune")
    dc = gpu_basis(shells).direct_contractor()
    D = torch.tensor(p[None], dtype=torch.float64, device="cuda:0").contiguous()
    sched = dc.autotune(D, threshold=1e-12, reps=1)
    assert sched and all(e["tuned"] and e["variant"] in (0, 1, 2, 3, 4) for e in sched)
    J, K = dc.jk([p], 1e-12)
    oJ, oK = O.Basis(shells).direct_jk([p], 1e-12)
    assert np.max(np.abs(J[0] - oJ[0])) < TOL and np.max(np.abs(K[0] - oK[0])) < TOL
    with pytest.raises(ig.IntegralError):
        dc.set_schedule([{"cls": (4, 4, 4, 4), "variant": 0}])
    with pytest.raises(ig.IntegralError):
        dc.set_schedule([{"cls": (4, 0, 0, 0), "variant": 0}])


def _support_density(shells, keep, seed):
    """Symmetric density whose nonzero blocks only touch the shells in `keep`:
    most quartets then fail the density mask, interleaved with ones that pass
    inside every wavefront's range."""
    n = B.nfunc(shells)
    off, o = [], 0
    for s in shells:
        off.append(o)
        o += (2 * s[0] + 1) if s[4] else (s[0] + 1) * (s[0] + 2) // 2
    rows = np.zeros(n, dtype=bool)
    for k in keep:
        nf = (2 * shells[k][0] + 1) if shells[k][4] else (shells[k][0] + 1) * (shells[k][0] + 2) // 2
        rows[off[k]:off[k] + nf] = True
    p = B.symmetric(n, seed)
    return p * np.outer(rows, rows)


@pytest.mark.parametrize("variant", [0, 1, 2, 3, 4])
def test_direct_jk_density_screened_lanes(variant):
    """Quartets dropped by the density mask next to kept ones in the same
    wavefront, across bra rows (regression: the wavefront-reduced bra J sum
    must be written at the kept lanes' bra pair)."""
    shells = B.grad_basis(True, max_l=2) + B.shared_centers_basis() + B.fd_basis()
    p = _support_density(shells, [0, 2, 7], 5)
    note("jk density-screened lanes variant %d" % variant)
    dc = gpu_basis(shells).direct_contractor()
    dc.set_schedule([dict(e, variant=variant) for e in dc.schedule()])
    J, K = dc.jk([p], 1e-12)
    ob = O.Basis(shells)
    oJ, oK = ob.direct_jk([p], 1e-12)
    assert dc.last_stats()[0] == ob.last_stats[0]
    assert 0 < ob.last_stats[0]
    assert np.max(np.abs(J[0] - oJ[0])) < TOL
    assert np.max(np.abs(K[0] - oK[0])) < TOL


@pytest.mark.parametrize("variant", [None, 0, 1, 2, 3, 4])
def test_direct_jk_twelve_atom_molecule(variant):
    """The benchmark's molecule recipe at 12 atoms (spherical s,s,p,p,d,
    3/3/3/3/1 primitives, nao = 156), two densities, tau = 1e-12: every
    kernel at a size where each wavefront walks many quartets and rows."""
    shells, rng = synthetic.molecule(20260007, 12, 8.0, [(0, 3), (0, 3), (1, 3), (1, 3), (2, 1)],
                                     (0.08, 40.0), True)
    n = B.nfunc(shells)
    dens = [synthetic.density(rng, n, 6), synthetic.density(rng, n, 6)]
    note("jk 12-atom molecule nao=%d variant %s" % (n, variant))
    dc = gpu_basis(shells).direct_contractor()
    if variant is not None:
        dc.set_schedule([dict(e, variant=variant) for e in dc.schedule()])
    J, K = dc.jk(dens, 1e-12)
    ob = O.Basis(shells)
    oJ, oK = ob.direct_jk(dens, 1e-12)
    assert dc.last_stats()[0] == ob.last_stats[0]
    for x in range(2):
        assert np.max(np.abs(J[x] - oJ[x])) < TOL
        assert np.max(np.abs(K[x] - oK[x])) < TOL


@pytest.mark.parametrize("variant", [None, 1, 2])
def test_mixed_precision(variant):
    """FP32 tables for quartets with Q_ij Q_kl below the cut.  At a 1e-8 cut
    |dJ|, |dK| <= 1e-9 against the oracle (the mixed-precision bar); at 1e-6
    the error grows with the FP32 share (<= 1e-7 here).  The screening counts
    do not change and FP64 is restored exactly."""
    shells, rng = synthetic.molecule(20260007, 12, 8.0, [(0, 3), (0, 3), (1, 3), (1, 3), (2, 1)],
                                     (0.08, 40.0), True)
    n = B.nfunc(shells)
    dens = [synthetic.density(rng, n, 6), synthetic.density(rng, n, 6)]
    note("mixed precision 12-atom variant %s" % variant)
    dc = gpu_basis(shells).direct_contractor()
    if variant is not None:
        dc.set_schedule([dict(e, variant=variant) for e in dc.schedule()])
    J64, K64 = dc.jk(dens, 1e-12)
    q64 = dc.last_stats()[0]
    oJ, oK = O.Basis(shells).direct_jk(dens, 1e-12)
    for cut, bar in ((1e-8, 1e-9), (1e-6, 1e-7)):
        dc.set_mixed_precision(cut)
        J, K = dc.jk(dens, 1e-12)
        assert dc.last_stats()[0] == q64
        err = max(max(np.max(np.abs(J[x] - oJ[x])), np.max(np.abs(K[x] - oK[x]))) for x in range(2))
        print("[case] cut %.0e max |dJ,dK| %.3e" % (cut, err), flush=True)
        assert err <= bar, (cut, err)
        # FP32 really ran: not bit-identical to FP64
        assert max(np.max(np.abs(J[x] - J64[x])) for x in range(2)) > 0.0
    dc.set_mixed_precision(0.0)
    J2, K2 = dc.jk(dens, 1e-12)
    assert np.max(np.abs(J2[0] - oJ[0])) < TOL and np.max(np.abs(K2[0] - oK[0])) < TOL


def test_row_sums_beyond_64k_lds():
    """The row-sum kernel with LDS accumulators over 64 KB (the 40-atom,
    two-density J/K benchmark molecule: (dp|dp) rows need about 69 KB) against
    the team kernel, which the oracle pins on the smaller molecules above."""
    import torch
    shells, dens, tau = synthetic.direct_jk_config()
    note("row sums with > 64 KB LDS, nao 520")
    dc = gpu_basis(shells).direct_contractor()
    D = torch.tensor(np.stack(dens), dtype=torch.float64, device="cuda:0").contiguous()
    out = {}
    for v in (0, 2):
        dc.set_schedule([dict(e, variant=v) for e in dc.schedule()])
        J, K = torch.zeros_like(D), torch.zeros_like(D)
        dc.jk_device(D, J, K, threshold=tau)
        torch.cuda.synchronize()
        out[v] = (J.cpu().numpy(), K.cpu().numpy())
    assert np.max(np.abs(out[0][0] - out[2][0])) < TOL
    assert np.max(np.abs(out[0][1] - out[2][1])) < TOL
