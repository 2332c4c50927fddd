"""Golden contracted Cartesian ERI blocks by McMurchie-Davidson in mpmath.

An algorithm independent of both reference engines (Hermite expansion of the
Gaussian products, Hermite Coulomb integrals R_tuv from the Boys function),
the same cross-check idea as crates/integral/tests/eri_cross_algorithm.rs.
Primitives carry the shell-level normalisation N(alpha; l,0,0) times the raw
coefficient (crates/integral/src/shell.rs:182-187); blocks are row-major
(a,b,c,d) in the Cartesian order lx desc, ly desc.  Writes md_eri.json.
"""
import json
import os

import mpmath as mp

mp.mp.dps = 40
HERE = os.path.dirname(os.path.abspath(__file__))

SHELLS = [  # (l, center, exponents, coefficients)
    (0, (0.0, 0.0, 0.0), [1.9, 0.45], [0.6, 0.5]),
    (1, (0.0, 0.0, 0.0), [0.8, 0.3], [0.7, 0.4]),
    (2, (0.9, -0.4, 0.6), [0.7], [1.0]),
    (0, (0.9, -0.4, 0.6), [0.5], [1.0]),
    (3, (-0.7, 1.1, -0.3), [0.6], [1.0]),
    (1, (-0.7, 1.1, -0.3), [1.1, 0.4], [0.3, 0.8]),
    (2, (0.2, 0.5, 1.3), [1.4, 0.35], [0.5, 0.7]),
]
QUARTETS = [(0, 0, 0, 0), (1, 0, 3, 3), (1, 1, 1, 1), (2, 3, 5, 0), (2, 2, 2, 2),
            (4, 0, 3, 1), (6, 2, 5, 1), (6, 6, 6, 6), (4, 5, 2, 6), (4, 4, 0, 0)]


def comps(l):
    return [(lx, ly, l - lx - ly) for lx in range(l, -1, -1) for ly in range(l - lx, -1, -1)]


def dfact(n):
    r = mp.mpf(1)
    while n > 1:
        r *= n
        n -= 2
    return r


def norm(a, l):
    return (2 * a / mp.pi) ** mp.mpf("0.75") * mp.sqrt((4 * a) ** l) / mp.sqrt(dfact(2 * l - 1))


def boys(n, T):
    if T == 0:
        return mp.mpf(1) / (2 * n + 1)
    a = mp.mpf(n) + mp.mpf(1) / 2
    return mp.gammainc(a, 0, T) / (2 * T ** a)


def hermite_E(i, j, a, b, XA, XB):
    p = a + b
    mu = a * b / p
    XAB = XA - XB
    XPA = (a * XA + b * XB) / p - XA
    XPB = (a * XA + b * XB) / p - XB
    E = {(0, 0, 0): mp.exp(-mu * XAB ** 2)}

    def get(ii, jj, t):
        if t < 0 or t > ii + jj:
            return mp.mpf(0)
        return E[(ii, jj, t)]

    for ii in range(i + 1):
        for jj in range(j + 1):
            if ii == 0 and jj == 0:
                continue
            for t in range(ii + jj + 1):
                if ii > 0:
                    v = get(ii - 1, jj, t - 1) / (2 * p) + XPA * get(ii - 1, jj, t) + (t + 1) * get(ii - 1, jj, t + 1)
                else:
                    v = get(ii, jj - 1, t - 1) / (2 * p) + XPB * get(ii, jj - 1, t) + (t + 1) * get(ii, jj - 1, t + 1)
                E[(ii, jj, t)] = v
    return lambda t: get(i, j, t)


def hermite_R(tmax, rho, PQ):
    T = rho * sum(x * x for x in PQ)
    nmax = tmax
    R = {}
    for n in range(nmax + 1):
        R[(0, 0, 0, n)] = (-2 * rho) ** n * boys(n, T)

    def r(t, u, v, n):
        if t < 0 or u < 0 or v < 0:
            return mp.mpf(0)
        key = (t, u, v, n)
        if key in R:
            return R[key]
        if t > 0:
            val = (t - 1) * r(t - 2, u, v, n + 1) + PQ[0] * r(t - 1, u, v, n + 1)
        elif u > 0:
            val = (u - 1) * r(t, u - 2, v, n + 1) + PQ[1] * r(t, u - 1, v, n + 1)
        else:
            val = (v - 1) * r(t, u, v - 2, n + 1) + PQ[2] * r(t, u, v - 1, n + 1)
        R[key] = val
        return val
    return lambda t, u, v: r(t, u, v, 0)


def prim_eri(la, a, A, lb, b, B, lc, c, C, ld, d, D):
    p, q = a + b, c + d
    P = [(a * A[k] + b * B[k]) / p for k in range(3)]
    Q = [(c * C[k] + d * D[k]) / q for k in range(3)]
    rho = p * q / (p + q)
    Rf = hermite_R(la + lb + lc + ld, rho, [P[k] - Q[k] for k in range(3)])
    pref = 2 * mp.pi ** mp.mpf("2.5") / (p * q * mp.sqrt(p + q))
    out = []
    for ca in comps(la):
        for cb in comps(lb):
            Eab = [hermite_E(ca[k], cb[k], a, b, A[k], B[k]) for k in range(3)]
            for cc in comps(lc):
                for cd in comps(ld):
                    Ecd = [hermite_E(cc[k], cd[k], c, d, C[k], D[k]) for k in range(3)]
                    s = mp.mpf(0)
                    for t in range(ca[0] + cb[0] + 1):
                        for u in range(ca[1] + cb[1] + 1):
                            for v in range(ca[2] + cb[2] + 1):
                                e1 = Eab[0](t) * Eab[1](u) * Eab[2](v)
                                for tt in range(cc[0] + cd[0] + 1):
                                    for uu in range(cc[1] + cd[1] + 1):
                                        for vv in range(cc[2] + cd[2] + 1):
                                            e2 = Ecd[0](tt) * Ecd[1](uu) * Ecd[2](vv)
                                            sgn = -1 if (tt + uu + vv) % 2 else 1
                                            s += e1 * e2 * sgn * Rf(t + tt, u + uu, v + vv)
                    out.append(pref * s)
    return out


def contracted(q):
    sh = [SHELLS[i] for i in q]
    total = None
    import itertools
    for prims in itertools.product(*[range(len(s[2])) for s in sh]):
        args, scale = [], mp.mpf(1)
        for s, k in zip(sh, prims):
            al = mp.mpf(s[2][k])
            scale *= mp.mpf(s[3][k]) * norm(al, s[0])
            args += [s[0], al, [mp.mpf(x) for x in s[1]]]
        blk = prim_eri(*args)
        blk = [scale * v for v in blk]
        total = blk if total is None else [x + y for x, y in zip(total, blk)]
    return [float(v) for v in total]


def main():
    out = {"shells": SHELLS, "quartets": []}
    for q in QUARTETS:
        out["quartets"].append({"shells": list(q), "block": contracted(q)})
        print(q, flush=True)
    with open(os.path.join(HERE, "md_eri.json"), "w") as f:
        json.dump(out, f)


if __name__ == "__main__":
    main()
