"""Algorithmic floating-point work of one canonical shell quartet.

F = min(F_Rys, F_HGP) + F_c2s + F_digest, counted per evaluated quartet with
K primitive quartets:

* F_Rys = K f(la,lb,lc,ld): Rys quadrature with n = L/2 + 1 roots;
  25 + 70n for the prefactor and roots/weights, per root 12 + 3 x (2D VRR and
  both transfers), and 4n N4 for the assembly (engine/rys.rs:121-308).
* F_HGP = K (30 header + (36 + 3L) Boys + VRR + 2 n_addr(ne) n_addr(nf)
  extraction) + HRR: Obara-Saika vertical recurrence per primitive quartet,
  3 flops per raise, +4 with the e_i/2p term, +2 with the cross term, over
  m <= L - deg(e) - deg(f); horizontal transfers once per contracted quartet,
  2 flops per entry (engine/os_eri.rs:26-41, 1062-1215, 3080-3140).
* F_c2s: per-quartet Cartesian->spherical transforms.  The J/K and gradient
  paths of this library transform the density and J/K matrices once per call
  instead (O(nao^2 l) flops), so their per-quartet F_c2s is 0; the dense
  tensor path transforms every spherical axis of the block (2 n_out n_in x
  the other dimensions per axis).
* F_digest: 4 flops per permutation image of every element per density
  (direct.rs:353-404), 32 N4_func per density for a quartet of four distinct
  shells.
"""


def ncart(l):
    return (l + 1) * (l + 2) // 2


def naddr(l):
    return (l + 1) * (l + 2) * (l + 3) // 6


def rys_prim(la, lb, lc, ld):
    n = (la + lb + lc + ld) // 2 + 1
    nb, nk = la + lb, lc + ld
    n4 = ncart(la) * ncart(lb) * ncart(lc) * ncart(ld)
    hb = sum(nb - j + 1 for j in range(1, lb + 1))
    hk = sum(nk - l + 1 for l in range(1, ld + 1))
    per_axis = 5 * (nb + 1) * nk + 3 * nb + 2 * (nk + 1) * hb + 2 * (la + 1) * (lb + 1) * hk
    return 25 + 70 * n + n * (12 + 3 * per_axis) + 4 * n * n4


def _comps(d):
    return [(lx, ly, d - lx - ly) for lx in range(d, -1, -1) for ly in range(d - lx, -1, -1)]


def hgp_prim(la, lb, lc, ld):
    ne, nf = la + lb, lc + ld
    L = ne + nf
    vrr = 0
    for de in range(1, ne + 1):  # bra raise, f = 0
        for e in _comps(de):
            i = 0 if e[0] else (1 if e[1] else 2)
            has2 = e[i] >= 2
            vrr += (L - de + 1) * (3 + (4 if has2 else 0))
    for df in range(1, nf + 1):  # ket raise
        for f in _comps(df):
            j = 0 if f[0] else (1 if f[1] else 2)
            has2 = f[j] >= 2
            for de in range(0, ne + 1):
                for e in _comps(de):
                    cross = e[j] > 0
                    vrr += (L - de - df + 1) * (3 + (4 if has2 else 0) + (2 if cross else 0))
    return 30 + 36 + 3 * L + vrr + 2 * naddr(ne) * naddr(nf)


def hgp_hrr(la, lb, lc, ld):
    ne, nf = la + lb, lc + ld
    nfu = naddr(nf) - naddr(lc - 1)
    bra = sum(ncart(da) * ncart(j) * nfu for j in range(1, lb + 1) for da in range(la, ne - j + 1))
    ket = ncart(la) * ncart(lb) * sum(ncart(dc) * ncart(k) for k in range(1, ld + 1)
                                      for dc in range(lc, nf - k + 1))
    return 2 * (bra + ket)


def c2s_flops(ls, spherical):
    dims = [ncart(l) for l in ls]
    total = 0
    for ax in range(4):
        if not spherical[ax]:
            continue
        nout = 2 * ls[ax] + 1
        other = 1
        for k in range(4):
            if k != ax:
                other *= dims[k]
        total += 2 * nout * dims[ax] * other
        dims[ax] = nout
    return total


def quartet_flops(ls, K, spherical, ndens=1, c2s_in_loop=False):
    """Algorithmic flops of one quartet class (la, lb, lc, ld) with K primitive
    quartets and `ndens` densities in its mask."""
    f_rys = K * rys_prim(*ls)
    f_hgp = K * hgp_prim(*ls) + hgp_hrr(*ls)
    nfunc = 1
    for l, s in zip(ls, spherical):
        nfunc *= (2 * l + 1) if s else ncart(l)
    f_c2s = c2s_flops(ls, spherical) if c2s_in_loop else 0
    return min(f_rys, f_hgp) + f_c2s + 32 * nfunc * ndens


def grad_quartet_flops(ls, K, nterms=2):
    """Algorithmic flops of one canonical quartet's contribution to the
    density-contracted ERI gradient (direct_grad.rs:394-620): the shifted
    blocks of the first three centres -- raised (l+1) and, for l > 0, lowered
    (l-1) -- each at min(F_Rys, F_HGP) with the same K primitive quartets
    (the fourth centre follows from translational invariance); 2 flops per
    weight x shifted-element product on each of the 3 axes; the weight block,
    about 6 flops per element and density term (the 4-term exchange
    symmetrisation, direct_grad.rs:490-547)."""
    la, lb, lc, ld = ls
    total = 0.0
    for pos in range(3):
        for dl in (1, -1):
            sh = list(ls)
            sh[pos] += dl
            if sh[pos] < 0:
                continue
            f_rys = K * rys_prim(*sh)
            f_hgp = K * hgp_prim(*sh) + hgp_hrr(*sh)
            n4 = ncart(sh[0]) * ncart(sh[1]) * ncart(sh[2]) * ncart(sh[3])
            total += min(f_rys, f_hgp) + 2 * 3 * n4
    n4 = ncart(la) * ncart(lb) * ncart(lc) * ncart(ld)
    return total + 6 * n4 * nterms
