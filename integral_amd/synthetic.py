"""Seeded synthetic molecules and densities for the benchmark workloads.

One generator, deterministic across platforms: splitmix64 drives uniform
doubles in [0, 1); normals come from Box-Muller.  Seeds are 20260000 + the
workload number.  Positions are uniform in a cube with rejection to keep a
minimum separation; exponents are log-uniform in the stated range and sorted
descending within a shell; raw coefficients are uniform in (0.1, 1);
densities are D = C C^T with C ~ N(0, sigma), symmetric and row-major.
"""
import math

import numpy as np

from . import Basis, Shell, ShellKind

_MASK = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed):
        self.state = seed & _MASK

    def next_u64(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & _MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK
        return z ^ (z >> 31)

    def uniform(self):
        return (self.next_u64() >> 11) * (1.0 / 9007199254740992.0)

    def normals(self, n):
        out = np.empty(n)
        for k in range(0, n, 2):
            u1 = max(self.uniform(), 1e-300)
            u2 = self.uniform()
            r = math.sqrt(-2.0 * math.log(u1))
            out[k] = r * math.cos(2.0 * math.pi * u2)
            if k + 1 < n:
                out[k + 1] = r * math.sin(2.0 * math.pi * u2)
        return out


def positions(rng, natom, side, min_sep=1.5, max_tries=100000):
    pts = []
    tries = 0
    while len(pts) < natom:
        tries += 1
        if tries > max_tries:
            raise RuntimeError("cannot place %d atoms in a %.1f bohr cube" % (natom, side))
        p = np.array([rng.uniform() * side for _ in range(3)])
        if all(np.linalg.norm(p - q) >= min_sep for q in pts):
            pts.append(p)
    return pts


def _exps(rng, n, lo, hi):
    e = [math.exp(math.log(lo) + rng.uniform() * (math.log(hi) - math.log(lo))) for _ in range(n)]
    return sorted(e, reverse=True)


def molecule(seed, natom, side, shells_per_atom, exp_range, spherical):
    """shells_per_atom: list of (l, nprim).  Returns shell tuples
    (l, center, exps, coefs, spherical) in atom-major order."""
    rng = SplitMix64(seed)
    pts = positions(rng, natom, side)
    out = []
    for p in pts:
        for l, npr in shells_per_atom:
            e = _exps(rng, npr, *exp_range)
            c = [0.1 + 0.9 * rng.uniform() for _ in range(npr)]
            out.append((l, tuple(float(x) for x in p), e, c, spherical))
    return out, rng


def density(rng, nao, k, sigma=0.1):
    C = rng.normals(nao * k).reshape(nao, k) * sigma
    return C @ C.T


def dense_eri_config():
    """10 atoms in an 8 bohr cube, Cartesian s,s,p shells with 3 primitives,
    exponents log-uniform in [0.1, 50]: nao = 50."""
    shells, rng = molecule(20260001, 10, 8.0, [(0, 3), (0, 3), (1, 3)], (0.1, 50.0), False)
    return shells, rng


def direct_jk_config():
    """40 atoms in a 14 bohr cube, spherical s,s,p,p (3 primitives) and d (1
    primitive) per atom, exponents log-uniform in [0.08, 40]: nao = 520; two
    densities D = C C^T with C nao x 20 ~ N(0, 0.1); Schwarz threshold 1e-12."""
    shells, rng = molecule(20260002, 40, 14.0, [(0, 3), (0, 3), (1, 3), (1, 3), (2, 1)],
                           (0.08, 40.0), True)
    nao = sum(2 * s[0] + 1 for s in shells)
    dens = [density(rng, nao, 20) for _ in range(2)]
    return shells, dens, 1e-12


def gradient_config():
    """30 atoms in a 12 bohr cube, spherical s,s,p,p,d with 3/3/2/2/1
    primitives, exponents log-uniform in [0.1, 30]: nao = 390; D = C C^T with
    C nao x 15 ~ N(0, 0.1); Coulomb factor 1.0, exchange factor -0.5."""
    shells, rng = molecule(20260004, 30, 12.0, [(0, 3), (0, 3), (1, 2), (1, 2), (2, 1)],
                           (0.1, 30.0), True)
    nao = sum(2 * s[0] + 1 for s in shells)
    D = density(rng, nao, 15)
    return shells, D


def high_l_config():
    """20 atoms in a 10 bohr cube, spherical s,p,d,f,g per atom plus an l = 6
    shell on every fourth atom, one primitive each, exponents log-uniform in
    [0.2, 10]: nao = 565; one density D = C C^T with C nao x 20 ~ N(0, 0.1).
    Rys rules up to 13 roots."""
    rng = SplitMix64(20260005)
    pts = positions(rng, 20, 10.0)
    shells = []
    for k, p in enumerate(pts):
        ls = [0, 1, 2, 3, 4] + ([6] if k % 4 == 0 else [])
        for l in ls:
            e = _exps(rng, 1, 0.2, 10.0)
            c = [0.1 + 0.9 * rng.uniform()]
            shells.append((l, tuple(float(x) for x in p), e, c, True))
    nao = sum(2 * s[0] + 1 for s in shells)
    return shells, density(rng, nao, 20)


def batch_config(nmol=64):
    """nmol molecules with uniform atom counts in [5, 60] at 1 atom per
    20 bohr^3 (cube side (20 N)^(1/3), min separation 1.5 bohr), spherical
    s,s,p,p,d,f per atom (f one primitive, the others 1-3), exponents
    log-uniform in [0.05, 60]; one density per molecule D = C C^T with
    C nao x 20 ~ N(0, 0.1)."""
    rng = SplitMix64(20260003)
    mols = []
    for _ in range(nmol):
        natom = 5 + int(rng.uniform() * 56)
        side = (20.0 * natom) ** (1.0 / 3.0)
        pts = positions(rng, natom, side)
        shells = []
        for p in pts:
            for l in (0, 0, 1, 1, 2, 3):
                npr = 1 if l == 3 else 1 + int(rng.uniform() * 3)
                e = _exps(rng, npr, 0.05, 60.0)
                c = [0.1 + 0.9 * rng.uniform() for _ in range(npr)]
                shells.append((l, tuple(float(x) for x in p), e, c, True))
        nao = sum(2 * s[0] + 1 for s in shells)
        mols.append((shells, density(rng, nao, 20)))
    return mols


def to_basis(shells):
    return Basis([Shell(l, c, e, k, ShellKind.Spherical if sph else ShellKind.Cartesian)
                  for (l, c, e, k, sph) in shells])
