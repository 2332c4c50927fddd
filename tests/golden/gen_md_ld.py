"""Golden contracted Cartesian ERI blocks for high angular momentum (l = 4..6)
and for the erf-attenuated operator, by McMurchie-Davidson in x86 extended
precision (numpy.longdouble, 64-bit mantissa) with Boys values from mpmath.

McMurchie-Davidson is independent of both reference engines (OS/HGP and Rys):
Gaussian products expanded in Hermite Gaussians (E coefficients), Hermite
Coulomb integrals R_tuv from the Boys function, block = E_bra R E_ket^T.
The attenuated operator erf(omega r)/r enters through rho -> rho s and a
factor sqrt(s), s = omega^2 / (rho + omega^2), in R and the prefactor: the
construction crates/integral/tests/erf_kernel.rs:46-230 uses for its own
McMurchie-Davidson reference.

Conventions follow the reference: primitives carry N(alpha; l,0,0) times the
raw coefficient (crates/integral/src/shell.rs:182-187), Cartesian order lx
descending then ly descending (math/am.rs:3-31), blocks row-major (a,b,c,d).

Before writing anything the generator re-derives the 40-digit mpmath blocks
of md_eri.json (l <= 3) and stops unless they agree to 1e-15 relative, so the
extended-precision arithmetic is itself checked.  Large blocks are stored as
a deterministic sample of elements (index, value).  Writes md_eri_high.json
and md_erf.json.
"""
import itertools
import json
import os
import sys
import time

import mpmath as mp
import numpy as np

LD = np.longdouble
HERE = os.path.dirname(os.path.abspath(__file__))
mp.mp.dps = 34

# (l, center, exponents, coefficients); the first seven are md_eri.json's
SHELLS = [
    (0, (0.0, 0.0, 0.0), [1.9, 0.45], [0.6, 0.5]),
    (1, (0.0, 0.0, 0.0), [0.8, 0.3], [0.7, 0.4]),
    (2, (0.9, -0.4, 0.6), [0.7], [1.0]),
    (0, (0.9, -0.4, 0.6), [0.5], [1.0]),
    (3, (-0.7, 1.1, -0.3), [0.6], [1.0]),
    (1, (-0.7, 1.1, -0.3), [1.1, 0.4], [0.3, 0.8]),
    (2, (0.2, 0.5, 1.3), [1.4, 0.35], [0.5, 0.7]),
    (4, (0.3, -0.2, 0.1), [0.9, 0.35], [0.6, 0.5]),   # 7: g, contracted
    (5, (-0.5, 0.8, 0.4), [0.7], [1.0]),              # 8: h
    (6, (0.6, 0.4, -0.7), [0.55], [1.0]),             # 9: i
    (4, (-0.4, -0.6, 0.9), [1.3], [1.0]),             # 10: g on another centre
]
HIGH = [(7, 0, 7, 3), (7, 7, 0, 3), (8, 1, 2, 0), (9, 0, 9, 3), (7, 10, 1, 0),
        (8, 8, 0, 0), (9, 7, 8, 2), (7, 7, 7, 7), (10, 9, 5, 1)]
ERF_OMEGAS = [0.4, 2.5]
ERF = [(0, 0, 0, 0), (1, 0, 3, 3), (1, 1, 1, 1), (2, 3, 5, 0), (2, 2, 2, 2), (4, 0, 3, 1),
       (6, 2, 5, 1), (6, 6, 6, 6), (4, 4, 0, 0), (7, 0, 7, 3), (8, 1, 2, 0)]
SAMPLE_MAX = 3000  # blocks with more elements keep a strided sample of about this many


def comps(l):
    return [(lx, ly, l - lx - ly) for lx in range(l, -1, -1) for ly in range(l - lx, -1, -1)]


def ld(x):
    return LD(mp.nstr(mp.mpf(x), 30))


def norm(a, l):
    a = mp.mpf(a)
    df = mp.mpf(1)
    n = 2 * l - 1
    while n > 1:
        df *= n
        n -= 2
    return (2 * a / mp.pi) ** mp.mpf("0.75") * mp.sqrt((4 * a) ** l) / mp.sqrt(df)


def boys(nmax, T):
    """F_0..F_nmax(T) (mpmath incomplete gamma), as longdouble."""
    out = []
    for n in range(nmax + 1):
        if T == 0:
            v = mp.mpf(1) / (2 * n + 1)
        else:
            a = mp.mpf(n) + mp.mpf(1) / 2
            v = mp.gammainc(a, 0, T) / (2 * T ** a)
        out.append(ld(v))
    return out


def e_table(la, lb, a, b, xa, xb):
    """E[i, j, t], i <= la, j <= lb, t <= i + j, for one axis."""
    p = a + b
    xp = (a * xa + b * xb) / p
    E = np.zeros((la + 1, lb + 1, la + lb + 2), dtype=LD)
    E[0, 0, 0] = np.exp(-(a * b / p) * (xa - xb) ** 2)
    h = LD(1) / (2 * p)
    for i in range(la + 1):
        for j in range(lb + 1):
            if i == 0 and j == 0:
                continue
            for t in range(i + j + 1):
                if i > 0:
                    prev, x = E[i - 1, j], xp - xa
                else:
                    prev, x = E[i, j - 1], xp - xb
                v = x * prev[t] + (t + 1) * prev[t + 1]
                if t > 0:
                    v += h * prev[t - 1]
                E[i, j, t] = v
    return E


def r_table(L, rho, PQ, T_mp):
    """R_tuv (n = 0) for t + u + v <= L from R^n_000 = (-2 rho)^n F_n(T)."""
    F = boys(L, T_mp)
    R = {}
    for n in range(L + 1):
        R[(0, 0, 0, n)] = (-2 * rho) ** n * F[n]

    def r(t, u, v, n):
        if t < 0 or u < 0 or v < 0:
            return LD(0)
        key = (t, u, v, n)
        if key not in R:
            if t > 0:
                val = (t - 1) * r(t - 2, u, v, n + 1) + PQ[0] * r(t - 1, u, v, n + 1)
            elif u > 0:
                val = (u - 1) * r(t, u - 2, v, n + 1) + PQ[1] * r(t, u - 1, v, n + 1)
            else:
                val = (v - 1) * r(t, u, v - 2, n + 1) + PQ[2] * r(t, u, v - 1, n + 1)
            R[key] = val
        return R[key]

    out = np.zeros((L + 1, L + 1, L + 1), dtype=LD)
    for t in range(L + 1):
        for u in range(L + 1 - t):
            for v in range(L + 1 - t - u):
                out[t, u, v] = r(t, u, v, 0)
    return out


def hermite_rows(la, lb, E, sign):
    """rows (a, b) x Hermite index (t, u, v) of the product E_x E_y E_z."""
    n = la + lb + 1
    rows = []
    for ca in comps(la):
        for cb in comps(lb):
            ex = E[0][ca[0], cb[0], :n]
            ey = E[1][ca[1], cb[1], :n]
            ez = E[2][ca[2], cb[2], :n]
            h = np.einsum("t,u,v->tuv", ex, ey, ez)
            if sign:
                s = np.array([(-1) ** k for k in range(n)], dtype=LD)
                h = h * np.einsum("t,u,v->tuv", s, s, s)
            rows.append(h.ravel())
    return np.array(rows, dtype=LD)


def prim_block(sh, alphas, omega):
    (la, A), (lb, B), (lc, C), (ld_, D) = sh
    a, b, c, d = (ld(x) for x in alphas)
    A, B, C, D = ([ld(x) for x in v] for v in (A, B, C, D))
    p, q = a + b, c + d
    P = [(a * A[k] + b * B[k]) / p for k in range(3)]
    Q = [(c * C[k] + d * D[k]) / q for k in range(3)]
    PQ = [P[k] - Q[k] for k in range(3)]
    rho = p * q / (p + q)
    pref = 2 * ld(mp.pi) ** LD(2.5) / (p * q * np.sqrt(p + q))
    # the same quantities in mpmath for the Boys argument
    am, bm, cm, dm = (mp.mpf(x) for x in alphas)
    pm, qm = am + bm, cm + dm
    Pm = [(am * mp.mpf(sh[0][1][k]) + bm * mp.mpf(sh[1][1][k])) / pm for k in range(3)]
    Qm = [(cm * mp.mpf(sh[2][1][k]) + dm * mp.mpf(sh[3][1][k])) / qm for k in range(3)]
    rhom = pm * qm / (pm + qm)
    if omega is not None:
        w2 = mp.mpf(omega) ** 2
        sm = w2 / (rhom + w2)
        rhom = rhom * sm
        rho = rho * ld(sm)
        pref = pref * ld(mp.sqrt(sm))
    Tm = rhom * sum((Pm[k] - Qm[k]) ** 2 for k in range(3))
    L = la + lb + lc + ld_
    R = r_table(L, rho, PQ, Tm)
    Eab = [e_table(la, lb, a, b, A[k], B[k]) for k in range(3)]
    Ecd = [e_table(lc, ld_, c, d, C[k], D[k]) for k in range(3)]
    bra = hermite_rows(la, lb, Eab, False)
    ket = hermite_rows(lc, ld_, Ecd, True)
    nb, nk = la + lb + 1, lc + ld_ + 1
    idx_b = np.array(list(itertools.product(range(nb), repeat=3)))
    idx_k = np.array(list(itertools.product(range(nk), repeat=3)))
    tt = idx_b[:, None, :] + idx_k[None, :, :]
    M = R[tt[..., 0], tt[..., 1], tt[..., 2]]
    return pref * (bra @ M @ ket.T)


def contracted(q, omega=None):
    sh = [SHELLS[i] for i in q]
    total = None
    for prims in itertools.product(*[range(len(s[2])) for s in sh]):
        scale = mp.mpf(1)
        for s, k in zip(sh, prims):
            scale *= mp.mpf(s[3][k]) * norm(s[2][k], s[0])
        blk = prim_block([(s[0], s[1]) for s in sh], [s[2][k] for s, k in zip(sh, prims)],
                         omega) * ld(scale)
        total = blk if total is None else total + blk
    return total.ravel()


def stored(block):
    n = block.size
    if n <= SAMPLE_MAX:
        idx = np.arange(n)
    else:
        idx = np.unique(np.concatenate([np.arange(0, n, n // SAMPLE_MAX + 1),
                                        [int(np.argmax(np.abs(block))), n - 1]]))
    return {"n": int(n), "index": [int(i) for i in idx], "value": [float(block[i]) for i in idx]}


def self_check():
    with open(os.path.join(HERE, "md_eri.json")) as f:
        g = json.load(f)
    worst = 0.0
    for rec in g["quartets"]:
        ref = np.array(rec["block"])
        v = contracted(tuple(rec["shells"])).astype(float)
        worst = max(worst, float(np.max(np.abs(v - ref)) / np.max(np.abs(ref))))
    print("extended precision vs 40-digit mpmath (l <= 3): worst rel %.2e" % worst, flush=True)
    if worst > 1e-15:
        sys.exit("extended-precision McMurchie-Davidson disagrees with md_eri.json")


def main():
    self_check()
    high = {"shells": SHELLS, "quartets": []}
    for q in HIGH:
        t0 = time.time()
        high["quartets"].append(dict(shells=list(q), **stored(contracted(q))))
        print("high", q, "%.1fs" % (time.time() - t0), flush=True)
    with open(os.path.join(HERE, "md_eri_high.json"), "w") as f:
        json.dump(high, f)
    erf = {"shells": SHELLS, "omegas": ERF_OMEGAS, "quartets": []}
    for w in ERF_OMEGAS:
        for q in ERF:
            erf["quartets"].append(dict(shells=list(q), omega=w, **stored(contracted(q, w))))
            print("erf", w, q, flush=True)
    with open(os.path.join(HERE, "md_erf.json"), "w") as f:
        json.dump(erf, f)


if __name__ == "__main__":
    main()
