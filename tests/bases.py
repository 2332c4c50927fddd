"""Test inputs shared by the CPU and GPU suites.

The small bases and density formulas restate the fixed inputs of the
reference's own tests (crates/integral/tests/direct_contractor.rs:3-45,
direct_grad_contractor.rs:32-120, 270-330, 420-470); shells are tuples
(l, center, exponents, coefficients, spherical).
"""
import numpy as np


def direct_contractor_basis(sph, max_l=3):
    """direct_contractor.rs:3-10 (s, p, d, f on four centres)."""
    shells = [
        (0, (0.0, 0.0, 0.0), [1.2, 0.35], [0.7, 0.4], sph),
        (1, (0.2, -0.1, 0.5), [0.8], [1.0], sph),
        (2, (-0.3, 0.4, -0.2), [0.6], [1.0], sph),
        (3, (0.1, 0.3, 0.7), [0.5], [1.0], sph),
    ]
    return [s for s in shells if s[0] <= max_l]


def direct_contractor_densities(n):
    """direct_contractor.rs:33-39: two non-symmetric densities."""
    return [np.array([(((p + 3 * shift) % 17) - 8.0) * 0.013 for p in range(n * n)]).reshape(n, n)
            for shift in range(2)]


def grad_basis(sph, max_l=3):
    """direct_grad_contractor.rs:32-44."""
    a, b, c = (0.0, 0.0, 0.0), (0.9, -0.4, 0.6), (-0.7, 1.1, -0.3)
    shells = [
        (0, a, [1.9, 0.45], [0.6, 0.5], sph),
        (1, a, [0.8, 0.3], [0.7, 0.4], sph),
        (2, b, [0.7], [1.0], sph),
        (0, b, [0.5], [1.0], sph),
        (3 if max_l >= 3 else 2, c, [0.6], [1.0], sph),
        (1, c, [1.1, 0.4], [0.3, 0.8], sph),
    ]
    return shells


def fd_basis():
    """direct_grad_contractor.rs:290-302 (mixed Cartesian/spherical)."""
    a, b, c = (0.0, 0.0, 0.0), (0.85, -0.3, 0.5), (-0.5, 0.9, 0.2)
    return [
        (0, a, [2.1, 0.5], [0.45, 0.6], False),
        (1, a, [0.9], [1.0], False),
        (2, b, [0.8], [1.0], True),
        (0, b, [0.6], [1.0], False),
        (1, c, [1.2, 0.4], [0.4, 0.7], False),
        (2, c, [0.65], [1.0], False),
    ]


def shared_centers_basis():
    """direct_grad_contractor.rs:420-434."""
    a, b = (0.1, -0.2, 0.3), (1.0, 0.7, -0.6)
    return [
        (0, a, [1.5, 0.4], [0.5, 0.6], False),
        (1, a, [0.8], [1.0], False),
        (2, a, [0.6], [1.0], True),
        (0, b, [0.7], [1.0], False),
        (1, b, [1.0, 0.35], [0.4, 0.7], False),
    ]


def symmetric(n, seed):
    """direct_grad_contractor.rs:46-56."""
    m = np.zeros((n, n))
    for p in range(n):
        for q in range(p + 1):
            v = (((p * 7 + q * 13 + seed * 5) % 23) - 11.0) * 0.017
            m[p, q] = v
            m[q, p] = v
    return m


def grad_terms(n):
    """direct_grad_contractor.rs:66-120: two Coulomb-type and two exchange-type
    terms, left != right in one term of each list."""
    p, da, db = symmetric(n, 1), symmetric(n, 2), symmetric(n, 3)
    coulomb = [(0.5, p, p), (0.3, da, db)]
    exchange = [(-0.5, da, da), (-0.2, db, p)]
    return coulomb, exchange


def dense_gamma(n, coulomb, exchange):
    g = np.zeros((n, n, n, n))
    for f, L, R in coulomb:
        g += f * np.einsum("pq,rs->pqrs", L, R)
    for f, L, R in exchange:
        g += f * np.einsum("pr,qs->pqrs", L, R)
    return g


def nfunc(shells):
    return sum((2 * s[0] + 1) if s[4] else (s[0] + 1) * (s[0] + 2) // 2 for s in shells)
