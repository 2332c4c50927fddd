"""The no-allocation J/K entry (ig_direct_jk_device_noalloc) against the
allocating device entry it mirrors (ig_direct_jk_device_ws), through the C
ABI: two calls in a row -- the nao 520 direct J/K workload with two
densities, then a smaller molecule with one density -- each equal to the
allocating entry's result to 1e-12; a call needing more than the fixed
array's 64 MiB is refused with IG_ERR_UNSUPPORTED before any device work."""
import numpy as np
import pytest
import torch

import integral_amd as ig
from integral_amd import synthetic
from tests import bases as B

pytestmark = pytest.mark.gpu

TOL = 1e-12


def note(msg):
    print("[case] " + msg, flush=True)


def dev(a):
    return torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda:0")


def compare(dc, d, tau):
    J0, K0 = torch.zeros_like(d), torch.zeros_like(d)
    dc.jk_device(d, J0, K0, threshold=tau, symmetric=True)
    J1, K1 = torch.zeros_like(d), torch.zeros_like(d)
    dc.jk_device_noalloc(d, J1, K1, threshold=tau, symmetric=True)
    torch.cuda.synchronize()
    ej = float((J0 - J1).abs().max())
    ek = float((K0 - K1).abs().max())
    note("noalloc vs allocating: max|dJ| %.2e max|dK| %.2e" % (ej, ek))
    assert ej < TOL and ek < TOL
    assert float(J0.abs().max()) > 0.0 and float(K0.abs().max()) > 0.0


def test_noalloc_matches_allocating_entry_twice_second_smaller():
    shells, dens, tau = synthetic.direct_jk_config()
    note("noalloc call 1: nao 520, 2 densities")
    dc = synthetic.to_basis(shells).direct_contractor()
    compare(dc, dev(np.stack(dens)), tau)
    small, rng = synthetic.molecule(20260203, 6, 6.0, [(0, 3), (0, 2), (1, 2), (1, 1), (2, 1)],
                                    (0.08, 40.0), True)
    n = B.nfunc(small)
    note("noalloc call 2: nao %d, 1 density" % n)
    dc2 = synthetic.to_basis(small).direct_contractor()
    compare(dc2, dev(synthetic.density(rng, n, 8)[None]), 1e-12)


def test_noalloc_refuses_calls_beyond_the_fixed_array():
    shells, dens, tau = synthetic.direct_jk_config()
    dc = synthetic.to_basis(shells).direct_contractor()
    n = dc.nao()
    d = torch.zeros(64, n, n, dtype=torch.float64, device="cuda:0")
    J, K = torch.zeros_like(d), torch.zeros_like(d)
    note("noalloc refusal: 64 densities at nao 520")
    with pytest.raises(ig.UnsupportedError):
        dc.jk_device_noalloc(d, J, K, threshold=tau)
    assert float(J.abs().max()) == 0.0
