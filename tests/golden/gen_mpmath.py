"""Golden vectors for the Boys function and the Rys quadrature from mpmath.

Boys: F_m(T) = gamma_lower(m + 1/2, T) / (2 T^(m + 1/2)), F_m(0) = 1/(2m+1),
at 150 significant digits.  Rys: the n-point Gauss rule of
d mu = 1/2 x^(-1/2) e^(-T x) dx on [0, 1], whose moments are F_k(T), from the
Chebyshev algorithm on the exact moments at 150 digits and the eigen-
decomposition of the Jacobi matrix (Golub-Welsch) in the same precision.  The
T grids include both sides of the branch seam 33 + 2.5 (2n - 1) of the
reference (crates/integral/src/math/rys.rs:174-185).  Writes
boys_mpmath.json and rys_mpmath.json next to this script.
"""
import json
import os

import mpmath as mp

mp.mp.dps = 150
HERE = os.path.dirname(os.path.abspath(__file__))


def boys(m, T):
    T = mp.mpf(T)
    if T == 0:
        return mp.mpf(1) / (2 * m + 1)
    a = mp.mpf(m) + mp.mpf(1) / 2
    return mp.gammainc(a, 0, T) / (2 * T ** a)


def rys(n, T):
    mu = [boys(k, T) for k in range(2 * n)]
    alpha, beta = [mp.mpf(0)] * n, [mp.mpf(0)] * n
    sig_prev = [mp.mpf(0)] * (2 * n)
    sig = list(mu)
    alpha[0] = mu[1] / mu[0]
    beta[0] = mu[0]
    for k in range(1, n):
        new = [mp.mpf(0)] * (2 * n)
        for l in range(k, 2 * n - k):
            new[l] = sig[l + 1] - alpha[k - 1] * sig[l] - beta[k - 1] * sig_prev[l]
        alpha[k] = new[k + 1] / new[k] - sig[k] / sig[k - 1]
        beta[k] = new[k] / sig[k - 1]
        sig_prev, sig = sig, new
    J = mp.matrix(n, n)
    for i in range(n):
        J[i, i] = alpha[i]
        if i + 1 < n:
            J[i, i + 1] = J[i + 1, i] = mp.sqrt(beta[i + 1])
    E, Q = mp.eigsy(J)
    pairs = sorted((E[i], beta[0] * Q[0, i] ** 2) for i in range(n))
    return [float(x) for x, _ in pairs], [float(w) for _, w in pairs]


def main():
    ts = [0.0, 1e-9, 0.03, 0.5, 1.0, 2.5, 7.3, 15.0, 29.9, 33.0, 40.0, 60.0, 92.9, 93.0, 120.0,
          500.0, 1e4, 1e8]
    out = []
    for m in list(range(0, 33)):
        for t in ts:
            out.append([m, t, float(boys(m, t))])
    with open(os.path.join(HERE, "boys_mpmath.json"), "w") as f:
        json.dump({"columns": ["m", "T", "F_m(T)"], "rows": out}, f)
    rows = []
    for n in range(1, 14):
        seam = 33.0 + 2.5 * (2 * n - 1)
        for t in [0.0, 0.3, 2.0, 7.5, 19.0, seam * 0.6, seam - 0.01, seam + 0.01, seam * 1.5]:
            x, w = rys(n, t)
            rows.append({"n": n, "T": t, "roots": x, "weights": w})
    with open(os.path.join(HERE, "rys_mpmath.json"), "w") as f:
        json.dump({"rows": rows}, f)
    print(len(out), len(rows))


if __name__ == "__main__":
    main()
