"""The C-ABI boundary on the GPU: per-worker workspaces (direct.rs:19-33,
direct_grad.rs:51-67), device-side ordering of calls that share scratch
across streams, the fully asynchronous device entry, the panic categories of
the reference's direct contractors as status codes (direct.rs:141-159,
direct_grad.rs:337-346, integrals.rs:316-321), and an allocation failure
that the library reports and survives."""
import ctypes

import numpy as np
import pytest
import torch

import integral_amd as ig
from integral_amd import synthetic
from oracle import oracle as O
from tests import bases as B

pytestmark = pytest.mark.gpu

TOL = 1e-12


def note(msg):
    print("[case] " + msg, flush=True)


def dev(a):
    return torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda:0")


def medium_molecule():
    """12 atoms of the s,s,p,p,d recipe of the direct J/K workload (nao 156):
    long enough per call that two calls on two streams overlap."""
    shells, rng = synthetic.molecule(20260202, 12, 8.0, [(0, 3), (0, 3), (1, 3), (1, 3), (2, 1)],
                                     (0.08, 40.0), True)
    n = B.nfunc(shells)
    return shells, [synthetic.density(rng, n, 20) for _ in range(2)]


def test_two_streams_on_one_contractor_are_ordered():
    """Calls on two streams share the contractor's scratch: the second waits
    for the first on the device, so both results are exact."""
    shells, (D1, D2) = medium_molecule()
    n = D1.shape[0]
    dc = synthetic.to_basis(shells).direct_contractor()
    ref = []
    for D in (D1, D2):
        J, K = torch.zeros(1, n, n, dtype=torch.float64, device="cuda"), None
        K = torch.zeros_like(J)
        dc.jk_device(dev(D[None]), J, K, threshold=1e-12)
        torch.cuda.synchronize()
        ref.append((J.cpu().numpy(), K.cpu().numpy()))
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    d1, d2 = dev(D1[None]), dev(D2[None])
    out = [(torch.zeros_like(d1), torch.zeros_like(d1)) for _ in range(2)]
    torch.cuda.synchronize()
    for rep in range(3):
        note("two streams, one contractor, rep %d" % rep)
        for J, K in out:
            J.zero_()
            K.zero_()
        torch.cuda.synchronize()
        for (d, (J, K), s) in ((d1, out[0], s1), (d2, out[1], s2)):
            with torch.cuda.stream(s):
                dc.jk_device(d, J, K, threshold=1e-12, stream=s, symmetric=True)
        torch.cuda.synchronize()
        for (J, K), (rj, rk) in zip(out, ref):
            assert np.max(np.abs(J.cpu().numpy() - rj)) < TOL
            assert np.max(np.abs(K.cpu().numpy() - rk)) < TOL


def test_workspaces_run_concurrently_and_match():
    shells, (D1, D2) = medium_molecule()
    n = D1.shape[0]
    dc = synthetic.to_basis(shells).direct_contractor()
    dens = [dev(D1[None]), dev(D2[None])]
    ref = []
    for d in dens:
        J, K = torch.zeros_like(d), torch.zeros_like(d)
        dc.jk_device(d, J, K, threshold=1e-12)
        torch.cuda.synchronize()
        ref.append((J.cpu().numpy(), K.cpu().numpy(), dc.last_stats()))
    ws = [ig.DirectWorkspace(), ig.DirectWorkspace()]
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    out = [(torch.zeros_like(d), torch.zeros_like(d)) for d in dens]
    torch.cuda.synchronize()
    note("two workspaces on two streams")
    for w, s, d, (J, K) in zip(ws, streams, dens, out):
        with torch.cuda.stream(s):
            dc.jk_device(d, J, K, threshold=1e-12, stream=s, workspace=w, symmetric=True)
    torch.cuda.synchronize()
    for w, (J, K), (rj, rk, st) in zip(ws, out, ref):
        assert np.max(np.abs(J.cpu().numpy() - rj)) < TOL
        assert np.max(np.abs(K.cpu().numpy() - rk)) < TOL
        assert w.last_stats() == st
    # a workspace follows the contractor's precision policy, read per call
    dc.set_mixed_precision(1e-8)
    J, K = torch.zeros_like(dens[0]), torch.zeros_like(dens[0])
    dc.jk_device(dens[0], J, K, threshold=1e-12, workspace=ws[0], symmetric=True)
    torch.cuda.synchronize()
    err = max(np.max(np.abs(J.cpu().numpy() - ref[0][0])), np.max(np.abs(K.cpu().numpy() - ref[0][1])))
    assert 0.0 < err < 1e-7
    dc.set_mixed_precision(0.0)


def test_async_entry_general_and_symmetric_paths_match_oracle():
    shells = B.direct_contractor_basis(True, max_l=2)
    n = B.nfunc(shells)
    dens = B.direct_contractor_densities(n)  # non-symmetric (direct_contractor.rs:33-39)
    oJ, oK = O.Basis(shells).direct_jk(dens, 0.0)
    dc = synthetic.to_basis(shells).direct_contractor()
    d = dev(np.stack(dens))
    for w in (None, ig.DirectWorkspace()):
        note("asynchronous entry, general path, workspace=%s" % (w is not None))
        J, K = torch.zeros_like(d), torch.zeros_like(d)
        dc.jk_device(d, J, K, threshold=0.0, workspace=w, symmetric=False)
        torch.cuda.synchronize()
        for k in range(2):
            assert np.max(np.abs(J[k].cpu().numpy().ravel() - oJ[k])) < TOL
            assert np.max(np.abs(K[k].cpu().numpy().ravel() - oK[k])) < TOL
    sym = [0.5 * (m + m.T) for m in dens]
    sJ, sK = O.Basis(shells).direct_jk(sym, 0.0)
    d = dev(np.stack(sym))
    J, K = torch.zeros_like(d), torch.zeros_like(d)
    dc.jk_device(d, J, K, threshold=0.0, symmetric=True)
    torch.cuda.synchronize()
    for k in range(2):
        assert np.max(np.abs(J[k].cpu().numpy().ravel() - sJ[k])) < TOL
        assert np.max(np.abs(K[k].cpu().numpy().ravel() - sK[k])) < TOL


def test_gradient_workspace_matches_contractor():
    shells = B.shared_centers_basis()
    n = B.nfunc(shells)
    coul, exch = B.grad_terms(n)
    gc = synthetic.to_basis(shells).direct_grad_contractor()
    dens = gc.density([ig.GradDensityTerm(*t) for t in coul], [ig.GradDensityTerm(*t) for t in exch])
    ref = gc.contract(dens, threshold=0.0)
    w = ig.DirectGradWorkspace()
    note("gradient workspace")
    assert np.max(np.abs(gc.contract(dens, threshold=0.0, workspace=w) - ref)) < TOL
    acc = np.zeros_like(ref)
    for bra in gc.bra_pairs():
        gc.accumulate_bra_pair(w, bra, dens, None, 0.0, acc)
    assert np.max(np.abs(acc - ref)) < TOL


def test_panic_categories_as_status_codes():
    """The reference's assert! panics of the direct contractors, each a status
    code through the C ABI, and the contractor still works afterwards."""
    lib = ig.library()
    shells = B.grad_basis(True, max_l=2)
    n = B.nfunc(shells)
    basis = synthetic.to_basis(shells)
    dc = basis.direct_contractor()
    D = B.symmetric(n, 1)
    J, K = np.zeros(n * n), np.zeros(n * n)
    ds = ig._ptrs([D.ravel()])
    js, ks = ig._ptrs([J]), ig._ptrs([K])
    bnd = np.zeros(len(shells) ** 2)
    note("panic categories: J/K")
    # non-canonical bra (direct.rs:149-159 takes bras from bra_pairs(): i >= j)
    assert lib.ig_direct_accumulate_bra_pair(dc._h, 0, 1, ds, ig._ptr(bnd), 1, 0.0, 1, js, ks,
                                             None) == ig.E_NONCANON
    assert lib.ig_direct_accumulate_bra_pair(dc._h, len(shells), 0, ds, ig._ptr(bnd), 1, 0.0, 1,
                                             js, ks, None) == ig.E_NONCANON
    # no densities / more than 64 (direct.rs:141-148)
    assert lib.ig_direct_jk(dc._h, ds, 0, 0.0, 1, js, ks, 0, -1, None) == ig.E_NODENS
    many = ig._ptrs([D.ravel()] * 65)
    assert lib.ig_direct_jk(dc._h, many, 65, 0.0, 1, ig._ptrs([J] * 65), ig._ptrs([K] * 65), 0, -1,
                            None) == ig.E_TOOMANY
    # bounds length (assert_eq! on lengths)
    out = np.zeros(len(shells) ** 2 - 1)
    assert lib.ig_direct_density_shell_bounds(dc._h, ds, 1, ig._ptr(out), out.size) == ig.E_LEN
    # a workspace of another contractor
    other = basis.direct_contractor()
    w = ig.DirectWorkspace()
    wh = w._handle(other)
    assert lib.ig_direct_jk_ws(dc._h, wh, ds, 1, 0.0, 1, js, ks, 0, -1, None) == ig.E_ARG
    note("panic categories: gradient")
    gc = basis.direct_grad_contractor()
    gc2 = synthetic.to_basis(shells).direct_grad_contractor()
    coul, exch = B.grad_terms(n)
    dens2 = gc2.density([ig.GradDensityTerm(*t) for t in coul], [])
    g = np.zeros(3 * 3)
    # density from a plan with another layout (direct_grad.rs:337-340)
    assert lib.ig_grad_contract(gc._h, dens2._h, 0, 0.0, 0.0, ig._ptr(g), g.size,
                                None) == ig.E_FOREIGN
    dens = gc.density([ig.GradDensityTerm(*t) for t in coul], [])
    # wrong out length (direct_grad.rs:341-343), omega <= 0 (integrals.rs:316-321),
    # non-canonical bra (direct_grad.rs:344-346)
    assert lib.ig_grad_contract(gc._h, dens._h, 0, 0.0, 0.0, ig._ptr(g), g.size - 1,
                                None) == ig.E_LEN
    assert lib.ig_grad_contract(gc._h, dens._h, 1, 0.0, 0.0, ig._ptr(g), g.size, None) == ig.E_OMEGA
    assert lib.ig_grad_contract(gc._h, dens._h, 1, float("inf"), 0.0, ig._ptr(g), g.size,
                                None) == ig.E_OMEGA
    assert lib.ig_grad_accumulate_bra_pair(gc._h, dens._h, 1, 2, 0, 0.0, 0.0, ig._ptr(g), g.size,
                                           None) == ig.E_NONCANON
    # everything still works
    Jg, Kg = dc.jk([D], 0.0)
    oJ, oK = O.Basis(shells).direct_jk([D], 0.0)
    assert np.max(np.abs(Jg[0] - oJ[0])) < TOL and np.max(np.abs(Kg[0] - oK[0])) < TOL


def test_allocation_failure_is_reported_and_survived():
    """A device allocation that cannot succeed (two nao_cart^4 buffers of
    about 1 TB each for a 600-function Cartesian basis) comes back as
    IG_ERR_DEVICE with the runtime's message; nothing is leaked into later
    calls.  The gamma pointer is never read: the buffers fail first."""
    lib = ig.library()
    shells = [(2, (1.7 * (k % 10), 1.7 * (k // 10), 0.0), [0.8], [1.0], False) for k in range(100)]
    basis = synthetic.to_basis(shells)
    assert basis.nao() == 600
    fake = ctypes.c_void_p(256)
    out = np.zeros(3 * 100)
    note("allocation failure: 2 x 1 TB")
    rc = lib.ig_eri_grad_contract_device(basis._h, 0, fake, 600 ** 4, 0, 0.0, ig._ptr(out),
                                         out.size, None)
    assert rc == ig.E_DEVICE
    assert b"hipMalloc" in lib.ig_last_error_message()
    small = B.shared_centers_basis()
    n = B.nfunc(small)
    gamma = np.random.default_rng(3).standard_normal((n, n, n, n)) * 0.1
    g = synthetic.to_basis(small).eri_grad_contract(gamma)
    o = O.Basis(small).eri_grad_contract(gamma)
    assert np.max(np.abs(g - o)) < TOL
