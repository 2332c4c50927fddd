"""Host walk of the batched J/K launches (ig_batch_check): the views and work
lists a batched graph would capture, for the six molecules of the batched
tests (nao 1 to 126, with and without f shells, Cartesian and spherical),
checked access by access on the host.  No batched kernel runs: the batched
path itself stays switched off (see test_gpu_batch.py)."""
import pytest

import integral_amd as ig
from integral_amd import synthetic
from tests.test_gpu_batch import batch_molecules

pytestmark = pytest.mark.gpu


def note(msg):
    print("[case] " + msg, flush=True)


@pytest.mark.parametrize("tau,want_k,fp32_below", [(1e-12, True, 0.0), (1e-12, False, 0.0),
                                                   (1e-9, True, 0.0), (1e-12, True, 1e-8)])
def test_batched_launch_indexing_in_range(tau, want_k, fp32_below):
    mols = batch_molecules()
    bases = [synthetic.to_basis(s) for s, _ in mols]
    note("batch check: nao %s, tau %g, exchange %s, fp32 below %g"
         % ([b.nao() for b in bases], tau, want_k, fp32_below))
    bc = ig.BatchContractor(bases)
    bc.set_mixed_precision(fp32_below)
    assert bc.check(tau, want_k) == ""


def test_batch_check_rejects_bad_threshold():
    bc = ig.BatchContractor([synthetic.to_basis(batch_molecules()[2][0])])
    with pytest.raises(ig.IntegralError):
        bc.check(float("nan"))
