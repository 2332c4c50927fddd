"""ctypes front end of the oracle (test infrastructure only).

The oracle is a plain-C CPU restatement of the reference library's
two-electron hot path (crates/integral/src/{math,engine}/*, integrals.rs,
direct.rs, direct_grad.rs, grad.rs).  Only the tests, the smoke check and the
CPU-baseline leg of the benchmark load it, and only as the checker or the
reported baseline -- the product never does.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

ENGINE_AUTO, ENGINE_OS, ENGINE_RYS = 0, 1, 2

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_i64p = ctypes.POINTER(ctypes.c_int64)
_dpp = ctypes.POINTER(_dp)


def _load():
    if not os.path.exists(_LIB_PATH):
        raise RuntimeError("oracle/liboracle.so is not built (run `make -C oracle`)")
    lib = ctypes.CDLL(_LIB_PATH)
    lib.or_boys.argtypes = [ctypes.c_int, ctypes.c_double, _dp]
    lib.or_boys_reference.argtypes = [ctypes.c_int, ctypes.c_double, _dp]
    lib.or_rys.argtypes = [ctypes.c_int, ctypes.c_double, _dp, _dp]
    lib.or_rys_reference.argtypes = [ctypes.c_int, ctypes.c_double, _dp, _dp]
    lib.or_rys_table_nsub.argtypes = [ctypes.c_int]
    lib.or_rys_table_worst.argtypes = [ctypes.c_int]
    lib.or_rys_table_worst.restype = ctypes.c_double
    lib.or_c2s_matrix.argtypes = [ctypes.c_int, _dp]
    lib.or_cart_norm.argtypes = [ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    lib.or_cart_norm.restype = ctypes.c_double
    lib.or_basis_create.argtypes = [ctypes.c_int, _ip, _ip, _dp, _ip, _dp, _dp,
                                    ctypes.POINTER(ctypes.c_void_p)]
    lib.or_basis_destroy.argtypes = [ctypes.c_void_p]
    for f in ("or_basis_nao", "or_basis_nao_cart", "or_basis_natom", "or_basis_nsh"):
        getattr(lib, f).argtypes = [ctypes.c_void_p]
    lib.or_eri_block.argtypes = [ctypes.c_void_p, ctypes.c_int] + [ctypes.c_int] * 4 + [_dp]
    lib.or_eri_block_cart.argtypes = [ctypes.c_void_p, ctypes.c_int] + [ctypes.c_int] * 4 + [_dp]
    lib.or_eri_dense.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, ctypes.c_int]
    lib.or_schwarz.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, ctypes.c_int]
    lib.or_eri_screened.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, _dp, _i64p,
                                    ctypes.c_int]
    lib.or_density_shell_bounds.argtypes = [ctypes.c_void_p, ctypes.c_int, _dpp, _dp]
    lib.or_direct_jk.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, ctypes.c_int, _dpp, _dp,
                                 ctypes.c_double, ctypes.c_int, _dpp, _dpp, ctypes.c_int64,
                                 ctypes.c_int64, ctypes.c_int, _i64p]
    lib.or_direct_grad.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, ctypes.c_int, _dp, _dpp,
                                   _dpp, ctypes.c_int, _dp, _dpp, _dpp, ctypes.c_double, _dp,
                                   ctypes.c_int64, ctypes.c_int64, ctypes.c_int, _i64p]
    lib.or_eri_grad_contract.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, ctypes.c_int]
    lib.or_last_error.restype = ctypes.c_char_p
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def _ptr(a):
    return a.ctypes.data_as(_dp)


def _ptrs(arrs):
    return (_dp * len(arrs))(*[_ptr(a) for a in arrs]) if arrs else None


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def default_threads():
    return int(os.environ.get("ORACLE_THREADS", os.cpu_count() or 1))


def boys(m_max, t, reference=False):
    out = np.zeros(m_max + 1)
    (lib().or_boys_reference if reference else lib().or_boys)(m_max, float(t), _ptr(out))
    return out


def rys(n, t, reference=False):
    r, w = np.zeros(n), np.zeros(n)
    (lib().or_rys_reference if reference else lib().or_rys)(n, float(t), _ptr(r), _ptr(w))
    return r, w


def rys_table_info(n):
    return lib().or_rys_table_nsub(n), lib().or_rys_table_worst(n)


def c2s_matrix(l):
    out = np.zeros((2 * l + 1) * (l + 1) * (l + 2) // 2)
    lib().or_c2s_matrix(l, _ptr(out))
    return out.reshape(2 * l + 1, (l + 1) * (l + 2) // 2)


def cart_norm(alpha, lx, ly, lz):
    return lib().or_cart_norm(alpha, lx, ly, lz)


class Basis:
    """Oracle basis built from (l, center, exps, coefs, spherical) tuples."""

    def __init__(self, shells):
        n = len(shells)
        self.shells = shells
        ls = (ctypes.c_int * max(n, 1))(*[s[0] for s in shells])
        kinds = (ctypes.c_int * max(n, 1))(*[1 if s[4] else 0 for s in shells])
        centers = _f64([c for s in shells for c in s[1]] or [0.0])
        nprim = (ctypes.c_int * max(n, 1))(*[len(s[2]) for s in shells])
        exps = _f64([e for s in shells for e in s[2]] or [0.0])
        coefs = _f64([c for s in shells for c in s[3]] or [0.0])
        h = ctypes.c_void_p()
        rc = lib().or_basis_create(n, ls, kinds, _ptr(centers), nprim, _ptr(exps), _ptr(coefs),
                                   ctypes.byref(h))
        if rc != 0:
            raise ValueError("oracle basis: %s (code %d)" % (lib().or_last_error().decode(), rc))
        self._h = h
        self.nao = lib().or_basis_nao(h)
        self.nao_cart = lib().or_basis_nao_cart(h)
        self.natom = lib().or_basis_natom(h)
        self.nsh = lib().or_basis_nsh(h)

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.or_basis_destroy(self._h)
            self._h = None

    def nfunc(self, s):
        l, _, _, _, sph = self.shells[s]
        return 2 * l + 1 if sph else (l + 1) * (l + 2) // 2

    def eri_block(self, i, j, k, l, engine=ENGINE_AUTO, cartesian=False):
        if cartesian:
            dims = [(self.shells[s][0] + 1) * (self.shells[s][0] + 2) // 2 for s in (i, j, k, l)]
        else:
            dims = [self.nfunc(s) for s in (i, j, k, l)]
        out = np.zeros(int(np.prod(dims)))
        f = lib().or_eri_block_cart if cartesian else lib().or_eri_block
        f(self._h, engine, i, j, k, l, _ptr(out))
        return out.reshape(dims)

    def packed_ao_pairs(self):
        """Shell-pair-blocked AO pairs (p, q), p >= q (packed_eri.rs:99-110):
        canonical shell pairs (i >= j) row-major, inside one pair (a, b)
        row-major with b <= a when i == j.  Returns (pairs, pair_start)."""
        off = np.concatenate([[0], np.cumsum([self.nfunc(s) for s in range(self.nsh)])])
        pairs, start = [], []
        for i in range(self.nsh):
            for j in range(i + 1):
                start.append(len(pairs))
                for a in range(self.nfunc(i)):
                    for b in range(a + 1 if i == j else self.nfunc(j)):
                        pairs.append((int(off[i] + a), int(off[j] + b)))
        start.append(len(pairs))
        return pairs, start

    def packed_eri(self, threshold=0.0, engine=ENGINE_AUTO):
        """PackedEriBuilder::build (packed_eri.rs:173-230, store_block
        :300-330): every canonical quartet (ket pair <= bra pair) evaluated
        once and stored at t(t+1)/2 + u; with threshold > 0, quartets with
        Q_ij Q_kl < threshold are skipped (left zero)."""
        pairs, start = self.packed_ao_pairs()
        m = len(pairs)
        out = np.zeros(m * (m + 1) // 2)
        canon = [(i, j) for i in range(self.nsh) for j in range(i + 1)]
        q = None
        if threshold > 0.0:
            qm = self.schwarz(engine).reshape(self.nsh, self.nsh)
            q = [qm[i, j] for i, j in canon]
        for bra, (i, j) in enumerate(canon):
            for ket in range(bra + 1):
                k, l = canon[ket]
                if q is not None and q[bra] * q[ket] < threshold:
                    continue
                blk = self.eri_block(i, j, k, l, engine)
                na, nb, nc, nd = blk.shape
                t = start[bra]
                for a in range(na):
                    for b in range(a + 1 if i == j else nb):
                        u = start[ket]
                        for c in range(nc):
                            for d in range(c + 1 if k == l else nd):
                                if u > t:
                                    break
                                out[t * (t + 1) // 2 + u] = blk[a, b, c, d]
                                u += 1
                        t += 1
        return out

    def eri(self, engine=ENGINE_AUTO, threads=None):
        n = self.nao
        out = np.zeros(n ** 4)
        lib().or_eri_dense(self._h, engine, _ptr(out), threads or default_threads())
        return out

    def schwarz(self, engine=ENGINE_AUTO, threads=None):
        q = np.zeros(self.nsh * self.nsh)
        lib().or_schwarz(self._h, engine, _ptr(q), threads or default_threads())
        return q

    def eri_screened(self, tau, engine=ENGINE_AUTO, threads=None):
        out = np.zeros(self.nao ** 4)
        st = np.zeros(2, dtype=np.int64)
        lib().or_eri_screened(self._h, engine, float(tau), _ptr(out),
                              st.ctypes.data_as(_i64p), threads or default_threads())
        return out, (int(st[0]), int(st[1]))

    def density_shell_bounds(self, dens):
        dens = [_f64(d) for d in dens]
        out = np.zeros(len(dens) * self.nsh * self.nsh)
        lib().or_density_shell_bounds(self._h, len(dens), _ptrs(dens), _ptr(out))
        return out

    def direct_jk(self, dens, threshold=0.0, want_exchange=True, engine=ENGINE_AUTO,
                  schwarz=None, bounds=None, bra_range=(0, -1), threads=None):
        dens = [_f64(d).ravel() for d in dens]
        n2 = self.nao * self.nao
        J = [np.zeros(n2) for _ in dens]
        K = [np.zeros(n2) for _ in dens]
        if bounds is None:
            bounds = self.density_shell_bounds(dens)
        bounds = _f64(bounds)
        q = None if schwarz is None else _ptr(_f64(schwarz))
        st = np.zeros(2, dtype=np.int64)
        rc = lib().or_direct_jk(self._h, engine, q, len(dens), _ptrs(dens), _ptr(bounds),
                                float(threshold), 1 if want_exchange else 0, _ptrs(J), _ptrs(K),
                                bra_range[0], bra_range[1], threads or default_threads(),
                                st.ctypes.data_as(_i64p))
        if rc != 0:
            raise ValueError(lib().or_last_error().decode())
        self.last_stats = (int(st[0]), int(st[1]))
        return J, (K if want_exchange else None)

    def direct_grad(self, coulomb, exchange, threshold=0.0, engine=ENGINE_AUTO, schwarz=None,
                    bra_range=(0, -1), threads=None):
        """coulomb / exchange: lists of (factor, left, right)."""
        cf = _f64([t[0] for t in coulomb] or [0.0])
        cl = [_f64(t[1]).ravel() for t in coulomb]
        cr = [_f64(t[2]).ravel() for t in coulomb]
        xf = _f64([t[0] for t in exchange] or [0.0])
        xl = [_f64(t[1]).ravel() for t in exchange]
        xr = [_f64(t[2]).ravel() for t in exchange]
        out = np.zeros(self.natom * 3)
        q = None if schwarz is None else _ptr(_f64(schwarz))
        st = np.zeros(2, dtype=np.int64)
        rc = lib().or_direct_grad(self._h, engine, q, len(coulomb), _ptr(cf), _ptrs(cl),
                                  _ptrs(cr), len(exchange), _ptr(xf), _ptrs(xl), _ptrs(xr),
                                  float(threshold), _ptr(out), bra_range[0], bra_range[1],
                                  threads or default_threads(), st.ctypes.data_as(_i64p))
        if rc != 0:
            raise ValueError(lib().or_last_error().decode())
        self.last_stats = (int(st[0]), int(st[1]))
        return out.reshape(self.natom, 3)

    def eri_grad_contract(self, gamma, engine=ENGINE_AUTO, threads=None):
        out = np.zeros(self.natom * 3)
        rc = lib().or_eri_grad_contract(self._h, engine, _ptr(_f64(gamma).ravel()), _ptr(out),
                                        threads or default_threads())
        if rc != 0:
            raise ValueError(lib().or_last_error().decode())
        return out.reshape(self.natom, 3)


# ---------------------------------------------------------------------------
# density fitting (df.rs): 2- and 3-center Coulomb integrals through the
# 4-center blocks with a unit-s dummy on the aux index (df.rs:1-39)

def _unit_s(shell):
    """unit_s (df.rs:33-37): one zero-exponent primitive, coefficient 1."""
    return (0, tuple(shell[1]), [0.0], [1.0], False)


def df_eri_2c(aux, engine=ENGINE_AUTO):
    """Basis::eri_2c_with (df.rs:56-90): (P 1s|Q 1s) over canonical aux shell
    pairs, each evaluated once and mirrored; row-major [naux, naux]."""
    na = len(aux)
    ob = Basis(list(aux) + [_unit_s(s) for s in aux])
    nf = [ob.nfunc(p) for p in range(na)]
    off = np.concatenate([[0], np.cumsum(nf)]).astype(int)
    n = int(off[-1])
    out = np.zeros((n, n))
    for p in range(na):
        for q in range(p + 1):
            blk = ob.eri_block(p, na + p, q, na + q, engine).reshape(nf[p], nf[q])
            for a in range(nf[p]):
                for b in range(nf[q]):
                    out[off[p] + a, off[q] + b] = blk[a, b]
                    out[off[q] + b, off[p] + a] = blk[a, b]
    return out.ravel()


def df_eri_3c(main, aux, engine=ENGINE_AUTO):
    """Eri3cBuilder::build (df.rs:463-530): (ij|P 1s) for every canonical main
    pair and aux shell, written to the (i, j) band and, for i != j, the (j, i)
    band; row-major [nao, nao, naux], P fastest."""
    nm, na = len(main), len(aux)
    ob = Basis(list(main) + list(aux) + [_unit_s(s) for s in aux])
    nf = [ob.nfunc(s) for s in range(nm + na)]
    off = np.concatenate([[0], np.cumsum(nf[:nm])]).astype(int)
    aoff = np.concatenate([[0], np.cumsum(nf[nm:])]).astype(int)
    n, naux = int(off[-1]), int(aoff[-1])
    out = np.zeros((n, n, naux))
    for i in range(nm):
        for j in range(i + 1):
            for p in range(na):
                blk = ob.eri_block(i, j, nm + p, nm + na + p, engine)
                blk = blk.reshape(nf[i], nf[j], nf[nm + p])
                out[off[i]:off[i + 1], off[j]:off[j + 1], aoff[p]:aoff[p + 1]] = blk
                if i != j:
                    out[off[j]:off[j + 1], off[i]:off[i + 1], aoff[p]:aoff[p + 1]] = \
                        blk.transpose(1, 0, 2)
    return out.ravel()


def df_schwarz_aux(aux, engine=ENGINE_AUTO):
    """Basis::schwarz_aux_bounds_with (df.rs:257-280): sqrt(max |(P|P)|) per
    aux shell from the diagonal of its (P 1s|P 1s) block."""
    na = len(aux)
    ob = Basis(list(aux) + [_unit_s(s) for s in aux])
    out = np.zeros(na)
    for p in range(na):
        n = ob.nfunc(p)
        blk = ob.eri_block(p, na + p, p, na + p, engine).reshape(n, n)
        out[p] = np.sqrt(np.max(np.abs(np.diag(blk))))
    return out
