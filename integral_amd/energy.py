"""A differentiable two-electron energy: forward J/K, backward forces.

    E(R, D) = 1/2 * c_J * <D, J[D]> + 1/2 * c_K * <D, K[D]>
    J[D]_pq = sum_rs (pq|rs) D_rs,   K[D]_pq = sum_rs (pr|qs) D_rs

The forward pass is one direct J/K build on the GPU (`DirectContractor.jk`,
the contract of direct.rs:131-218). The backward pass is the adjoint
gradient kernel (`DirectGradContractor.contract`, direct_grad.rs:171-392)
with the Coulomb term (c_J/2, D, D) and the exchange term (c_K/2, D, D):

    dE/dR_A = sum_pqrs [c_J/2 D_pq D_rs + c_K/2 D_pr D_qs] d(pq|rs)/dR_A

so first-derivative integrals are contracted on the fly and nao^4 is never
formed. dE/dD = c_J J + c_K K comes from the forward pass for free.

The shells follow the atoms: a basis is a list of templates
(atom, l, exponents, coefficients, spherical) whose centres are taken from
the positions passed in, so the same templates can be re-evaluated at
displaced geometries (central finite differences).
"""
import numpy as np

from . import Basis, GradDensityTerm, Shell, ShellKind, UsageError


def templates_from_shells(shells):
    """Shell tuples (l, centre, exponents, coefficients, spherical) to
    templates plus the atom positions (atoms = distinct centres, in order)."""
    centres, index, out = [], {}, []
    for (l, c, e, k, sph) in shells:
        key = tuple(float(x) for x in c)
        if key not in index:
            index[key] = len(centres)
            centres.append(key)
        out.append((index[key], l, list(e), list(k), bool(sph)))
    return out, np.array(centres, dtype=np.float64)


def basis_at(templates, positions):
    pos = np.asarray(positions, dtype=np.float64)
    return Basis([Shell(l, tuple(pos[a]), e, k, ShellKind.Spherical if sph else ShellKind.Cartesian)
                  for (a, l, e, k, sph) in templates])


def _atom_rows(templates, basis):
    """Row of the gradient output (basis atoms, by first appearance of a
    centre) for each template atom index."""
    rows, first = {}, []
    for (a, *_rest) in templates:
        if a not in rows:
            rows[a] = len(first)
            first.append(a)
    if len(first) != len(basis.atoms()):
        raise UsageError("two atoms of the templates share a centre")
    return rows


def energy(templates, positions, D, coulomb=1.0, exchange=-0.5, threshold=0.0):
    """Forward pass: E and (J, K) at the given positions (one density)."""
    b = basis_at(templates, positions)
    D = np.ascontiguousarray(D, dtype=np.float64)
    want_k = exchange != 0.0
    J, K = b.direct_contractor().jk([D], threshold, want_exchange=want_k)
    n = b.nao()
    J = J[0].reshape(n, n)
    K = K[0].reshape(n, n) if want_k else np.zeros((n, n))
    E = 0.5 * coulomb * float(np.sum(D * J)) + 0.5 * exchange * float(np.sum(D * K))
    return E, J, K


def forces(templates, positions, D, coulomb=1.0, exchange=-0.5, threshold=0.0):
    """Backward pass: dE/dR (natom x 3, rows in template atom order)."""
    pos = np.asarray(positions, dtype=np.float64)
    b = basis_at(templates, pos)
    D = np.ascontiguousarray(D, dtype=np.float64)
    gc = b.direct_grad_contractor()
    cterms = [GradDensityTerm(0.5 * coulomb, D, D)] if coulomb != 0.0 else []
    xterms = [GradDensityTerm(0.5 * exchange, D, D)] if exchange != 0.0 else []
    g = gc.contract(gc.density(cterms, xterms), threshold=threshold)
    rows = _atom_rows(templates, b)
    out = np.zeros((pos.shape[0], 3))
    for a, r in rows.items():
        out[a] = g[r]
    return out


def finite_difference(templates, positions, D, coords, h=1e-4, **kw):
    """Central differences (E(R+h) - E(R-h)) / 2h for the (atom, axis) pairs
    in coords."""
    pos = np.asarray(positions, dtype=np.float64)
    out = []
    for a, x in coords:
        p, m = pos.copy(), pos.copy()
        p[a, x] += h
        m[a, x] -= h
        out.append((energy(templates, p, D, **kw)[0] - energy(templates, m, D, **kw)[0]) / (2 * h))
    return np.array(out)


try:
    import torch

    class JKEnergy(torch.autograd.Function):
        """torch.autograd wrapper: E = JKEnergy.apply(positions, D, templates,
        c_J, c_K, threshold); E.backward() fills positions.grad (the adjoint
        gradient kernel) and D.grad (c_J J + c_K K)."""

        @staticmethod
        def forward(ctx, positions, D, templates, coulomb=1.0, exchange=-0.5, threshold=0.0):
            pos = positions.detach().cpu().double().numpy()
            Dn = D.detach().cpu().double().numpy()
            E, J, K = energy(templates, pos, Dn, coulomb, exchange, threshold)
            ctx.save_for_backward(positions, D)
            ctx.args = (templates, coulomb, exchange, threshold)
            ctx.dD = coulomb * J + exchange * K
            return torch.tensor(E, dtype=torch.float64, device=positions.device)

        @staticmethod
        def backward(ctx, gE):
            positions, D = ctx.saved_tensors
            templates, coulomb, exchange, threshold = ctx.args
            g_pos = g_D = None
            if ctx.needs_input_grad[0]:
                f = forces(templates, positions.detach().cpu().double().numpy(),
                           D.detach().cpu().double().numpy(), coulomb, exchange, threshold)
                g_pos = gE * torch.as_tensor(f, dtype=positions.dtype, device=positions.device)
            if ctx.needs_input_grad[1]:
                g_D = gE * torch.as_tensor(ctx.dD, dtype=D.dtype, device=D.device)
            return g_pos, g_D, None, None, None, None
except ImportError:  # torch is optional for the numpy entry points above
    JKEnergy = None
