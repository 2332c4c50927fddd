"""The differentiable J/K energy (integral_amd.energy).

CPU: the convention itself, on the oracle -- forces from the Coulomb term
(c_J/2, D, D) and the exchange term (c_K/2, D, D) equal central finite
differences of E = c_J/2 <D, J[D]> + c_K/2 <D, K[D]>.
GPU: forward and backward through the HIP library against the oracle,
central differences at +-1e-4 bohr on the GPU energies, and D.grad."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import bases as B

CJ, CK = 1.0, -0.5


def small_system():
    shells = B.fd_basis()
    n = B.nfunc(shells)
    return shells, B.symmetric(n, 3)


def oracle_energy(shells, D):
    J, K = O.Basis(shells).direct_jk([D], 0.0)
    n = D.shape[0]
    return 0.5 * CJ * np.sum(D * J[0].reshape(n, n)) + 0.5 * CK * np.sum(D * K[0].reshape(n, n))


def displaced(shells, centre, axis, h):
    out = []
    for (l, c, e, k, sph) in shells:
        c = list(c)
        if tuple(c) == tuple(centre):
            c[axis] += h
        out.append((l, tuple(c), e, k, sph))
    return out


def test_energy_convention_on_oracle():
    shells, D = small_system()
    centres = []
    for s in shells:
        if s[1] not in centres:
            centres.append(s[1])
    g = O.Basis(shells).direct_grad([(0.5 * CJ, D, D)], [(0.5 * CK, D, D)], 0.0)
    h = 1e-4
    for a, c in enumerate(centres):
        for x in range(3):
            fd = (oracle_energy(displaced(shells, c, x, h), D)
                  - oracle_energy(displaced(shells, c, x, -h), D)) / (2 * h)
            assert abs(fd - g[a, x]) < 1e-7, (a, x, fd, g[a, x])


def test_templates_roundtrip():
    from integral_amd import energy as EN
    shells, _ = small_system()
    t, pos = EN.templates_from_shells(shells)
    assert pos.shape == (3, 3)
    assert [x[0] for x in t] == [0, 0, 1, 1, 2, 2]


@pytest.mark.gpu
def test_energy_forward_backward_on_gpu():
    import torch
    from integral_amd import energy as EN
    shells, D = small_system()
    print("[case] energy forward/backward fd_basis", flush=True)
    t, pos = EN.templates_from_shells(shells)
    P = torch.tensor(pos, dtype=torch.float64, requires_grad=True)
    Dt = torch.tensor(D, dtype=torch.float64, requires_grad=True)
    E = EN.JKEnergy.apply(P, Dt, t, CJ, CK, 0.0)
    assert abs(E.item() - oracle_energy(shells, D)) < 1e-12
    E.backward()
    og = O.Basis(shells).direct_grad([(0.5 * CJ, D, D)], [(0.5 * CK, D, D)], 0.0)
    assert np.max(np.abs(P.grad.numpy() - og)) < 1e-12
    J, K = O.Basis(shells).direct_jk([D], 0.0)
    n = D.shape[0]
    dD = CJ * J[0].reshape(n, n) + CK * K[0].reshape(n, n)
    assert np.max(np.abs(Dt.grad.numpy() - dD)) < 1e-12
    coords = [(a, x) for a in range(pos.shape[0]) for x in range(3)]
    fd = EN.finite_difference(t, pos, D, coords, h=1e-4, coulomb=CJ, exchange=CK)
    assert np.max(np.abs(fd - P.grad.numpy().ravel())) < 1e-7


@pytest.mark.gpu
def test_energy_gradient_config_fd():
    """30 atoms, s,s,p,p,d (nao = 390), Coulomb 1.0 / exchange -0.5: forces
    against central differences at +-1e-4 bohr on a few coordinates."""
    from integral_amd import energy as EN, synthetic
    shells, D = synthetic.gradient_config()
    print("[case] energy gradient config nao=390", flush=True)
    t, pos = EN.templates_from_shells(shells)
    f = EN.forces(t, pos, D, CJ, CK, 1e-14)
    assert np.max(np.abs(f.sum(axis=0))) < 1e-9
    coords = [(0, 0), (7, 1), (19, 2), (29, 0)]
    fd = EN.finite_difference(t, pos, D, coords, h=1e-4, coulomb=CJ, exchange=CK, threshold=1e-14)
    ref = np.array([f[a, x] for a, x in coords])
    assert np.max(np.abs(fd - ref)) < 1e-6, (fd, ref)
