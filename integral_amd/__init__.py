"""integral_amd -- MI355X (gfx950) implementation of the two-electron hot path
of the Rust library `integral` (v0.4.3), behind a C ABI (include/integral_hip.h).

This module mirrors the reference's public API for that path over ctypes:

    Shell / Basis                 crates/integral/src/shell.rs:49-275
    Basis.eri / eri_screened / schwarz_bounds
                                  crates/integral/src/integrals.rs:664-983
    DirectContractor              crates/integral/src/direct.rs:39-331
    DirectGradContractor          crates/integral/src/direct_grad.rs:144-650
    Engine, EriKernel             crates/integral/src/integrals.rs:304-338
    IntegralError                 crates/integral/src/lib.rs:83-123

Every call runs on the GPU through libintegral_hip.so; there is no CPU
fallback.  Importing works without a GPU (so the CPU test suite can check
the boundary), but any compute call then raises DeviceError.
"""
import ctypes
import math
import enum
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libintegral_hip.so")

__all__ = [
    "Engine", "ShellKind", "EriKernel", "Shell", "Basis", "DirectContractor", "DirectWorkspace",
    "DirectBuffers", "DirectGradContractor", "DirectGradWorkspace", "GradDensityTerm",
    "IntegralError", "EmptyContraction", "MismatchedContraction", "AngularMomentumTooHigh",
    "AngularMomentumTooHighForGradient", "DeviceError", "UnsupportedError", "MAX_L",
    "MAX_GRAD_L", "abi_version", "library",
]

MAX_L = 6
MAX_GRAD_L = 5

# status codes, include/integral_hip.h
OK, E_EMPTY, E_MISMATCH, E_LHIGH, E_LGRAD, E_GAMMA = 0, 1, 2, 3, 4, 5
E_NODENS, E_TOOMANY, E_LEN, E_NONCANON, E_FOREIGN, E_OMEGA, E_ARG, E_UNSUP, E_DEVICE = (
    10, 11, 12, 13, 14, 15, 16, 17, 20)


class Engine(enum.IntEnum):
    """integrals.rs:330-338."""
    Auto = 0
    OsHgp = 1
    Rys = 2


class ShellKind(enum.IntEnum):
    """shell.rs:17-27."""
    Cartesian = 0
    Spherical = 1


class EriKernel:
    """integrals.rs:304-312: Coulomb, or the long-range erf(omega r)/r."""

    def __init__(self, omega=None):
        self.omega = omega

    @classmethod
    def coulomb(cls):
        return cls(None)

    @classmethod
    def erf(cls, omega):
        return cls(float(omega))

    def _args(self):
        """(kernel id, omega) for the C ABI: IG_KERNEL_COULOMB=0, IG_KERNEL_ERF=1."""
        return (0, 0.0) if self.omega is None else (1, float(self.omega))


class IntegralError(Exception):
    """lib.rs:83-123 and the panics of the direct contractors."""
    code = None


class EmptyContraction(IntegralError):
    code = E_EMPTY


class MismatchedContraction(IntegralError):
    code = E_MISMATCH


class AngularMomentumTooHigh(IntegralError):
    code = E_LHIGH


class AngularMomentumTooHighForGradient(IntegralError):
    code = E_LGRAD


class GammaLengthMismatch(IntegralError):
    code = E_GAMMA


class DeviceError(IntegralError):
    code = E_DEVICE


class UnsupportedError(IntegralError):
    code = E_UNSUP


class UsageError(IntegralError, ValueError):
    """Misuse the reference rejects with a panic (assert!)."""


_ERRORS = {c.code: c for c in (EmptyContraction, MismatchedContraction, AngularMomentumTooHigh,
                                AngularMomentumTooHighForGradient, GammaLengthMismatch,
                                DeviceError, UnsupportedError)}

_dp = ctypes.POINTER(ctypes.c_double)
_dpp = ctypes.POINTER(_dp)
_vp = ctypes.c_void_p


class _ShellDesc(ctypes.Structure):
    _fields_ = [("l", ctypes.c_int32), ("kind", ctypes.c_int32), ("center", ctypes.c_double * 3),
                ("n_exponents", ctypes.c_int32), ("n_coefficients", ctypes.c_int32),
                ("exponents", _dp), ("coefficients", _dp)]


class _ClassProfile(ctypes.Structure):
    _fields_ = [("la", ctypes.c_int32), ("lb", ctypes.c_int32), ("lc", ctypes.c_int32),
                ("ld", ctypes.c_int32), ("fp32", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("quartets", ctypes.c_int64),
                ("prim_quartets", ctypes.c_int64), ("ms", ctypes.c_double)]


_lib = None


class _ScheduleEntry(ctypes.Structure):
    _fields_ = [("la", ctypes.c_int32), ("lb", ctypes.c_int32), ("lc", ctypes.c_int32),
                ("ld", ctypes.c_int32), ("variant", ctypes.c_int32), ("tuned", ctypes.c_int32),
                ("ms", ctypes.c_double)]


def library():
    """Load libintegral_hip.so (built by __graft_entry__.build or `make -C integral_amd/csrc`)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("integral_amd: %s is not built; run `make -C integral_amd/csrc`" % LIB_PATH)
    # PyTorch ships its own HIP runtime (same soname libamdhip64.so.7).  Loaded
    # after ours, it would start a second runtime that finds no device; loaded
    # first, the library binds to it and both share one runtime and one device.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(LIB_PATH)
    sz = ctypes.c_size_t
    i64 = ctypes.c_int64
    lib.integral_abi_version.restype = ctypes.c_uint32
    lib.ig_last_error_message.restype = ctypes.c_char_p
    lib.ig_basis_create.argtypes = [ctypes.POINTER(_ShellDesc), sz, ctypes.POINTER(_vp)]
    lib.ig_basis_destroy.argtypes = [_vp]
    for f in ("ig_basis_nao", "ig_basis_nao_cart", "ig_basis_natom", "ig_basis_nshell"):
        getattr(lib, f).argtypes = [_vp]
        getattr(lib, f).restype = sz
    lib.ig_eri_dense.argtypes = [_vp, ctypes.c_int, _dp, sz, _vp]
    lib.ig_eri_screened.argtypes = [_vp, ctypes.c_int, ctypes.c_double, _dp, sz,
                                    ctypes.POINTER(i64), _vp]
    lib.ig_schwarz_bounds.argtypes = [_vp, ctypes.c_int, _dp, sz, _vp]
    lib.ig_direct_create.argtypes = [_vp, ctypes.c_int, ctypes.POINTER(_vp)]
    lib.ig_direct_destroy.argtypes = [_vp]
    lib.ig_direct_nao.argtypes = [_vp]
    lib.ig_direct_nao.restype = sz
    lib.ig_direct_n_bra_pairs.argtypes = [_vp]
    lib.ig_direct_n_bra_pairs.restype = sz
    lib.ig_direct_bra_pairs.argtypes = [_vp, ctypes.POINTER(ctypes.c_int32), sz]
    lib.ig_direct_density_shell_bounds.argtypes = [_vp, _dpp, sz, _dp, sz]
    lib.ig_direct_accumulate_bra_pair.argtypes = [_vp, ctypes.c_int32, ctypes.c_int32, _dpp, _dp,
                                                  sz, ctypes.c_double, ctypes.c_int, _dpp, _dpp,
                                                  _vp]
    lib.ig_direct_jk.argtypes = [_vp, _dpp, sz, ctypes.c_double, ctypes.c_int, _dpp, _dpp, i64,
                                 i64, _vp]
    lib.ig_direct_jk_device.argtypes = [_vp, _vp, sz, ctypes.c_double, ctypes.c_int, _vp, _vp,
                                        _vp]
    lib.ig_direct_last_stats.argtypes = [_vp, ctypes.POINTER(i64)]
    lib.ig_direct_workspace_create.argtypes = [_vp, ctypes.POINTER(_vp)]
    lib.ig_direct_workspace_destroy.argtypes = [_vp]
    lib.ig_direct_workspace_destroy.restype = None
    lib.ig_direct_accumulate_bra_pair_ws.argtypes = [_vp, _vp, ctypes.c_int32, ctypes.c_int32,
                                                     _dpp, _dp, sz, ctypes.c_double, ctypes.c_int,
                                                     _dpp, _dpp, _vp]
    lib.ig_direct_jk_ws.argtypes = [_vp, _vp, _dpp, sz, ctypes.c_double, ctypes.c_int, _dpp, _dpp,
                                    i64, i64, _vp]
    lib.ig_direct_jk_device_ws.argtypes = [_vp, _vp, _vp, sz, ctypes.c_double, ctypes.c_int, _vp,
                                           _vp, ctypes.c_int, _vp]
    lib.ig_direct_workspace_last_stats.argtypes = [_vp, ctypes.POINTER(i64)]
    lib.ig_direct_jk_device_noalloc.argtypes = [_vp, _vp, sz, ctypes.c_double, ctypes.c_int, _vp,
                                                _vp, ctypes.c_int, _vp]
    lib.ig_direct_set_profiling.argtypes = [_vp, ctypes.c_int]
    lib.ig_direct_profile_read.argtypes = [_vp, ctypes.POINTER(_ClassProfile), sz,
                                           ctypes.POINTER(sz)]
    lib.ig_direct_set_mixed_precision.argtypes = [_vp, ctypes.c_double]
    lib.ig_boys_eval.argtypes = [ctypes.c_int, _dp, sz, _dp]
    lib.ig_rys_eval.argtypes = [ctypes.c_int, _dp, sz, _dp, _dp]
    lib.ig_direct_autotune.argtypes = [_vp, _vp, sz, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                       _vp]
    lib.ig_direct_schedule_read.argtypes = [_vp, ctypes.POINTER(_ScheduleEntry), sz,
                                            ctypes.POINTER(sz)]
    lib.ig_direct_schedule_write.argtypes = [_vp, ctypes.POINTER(_ScheduleEntry), sz]
    lib.ig_eri_builder_fill.argtypes = [_vp, ctypes.c_int, i64, i64, _dp, sz, _vp]
    lib.ig_eri_2c.argtypes = [_vp, ctypes.c_int, _dp, sz, _vp]
    lib.ig_eri_kernel.argtypes = [_vp, ctypes.c_int, ctypes.c_double, _dp, sz, _vp]
    lib.ig_eri_builder_fill_kernel.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                               i64, i64, _dp, sz, _vp]
    lib.ig_eri_2c_kernel.argtypes = [_vp, ctypes.c_int, ctypes.c_double, _dp, sz, _vp]
    lib.ig_eri_3c_fill_kernel.argtypes = [_vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                          i64, i64, _dp, sz, _vp]
    lib.ig_schwarz_aux_bounds.argtypes = [_vp, ctypes.c_int, _dp, sz, _vp]
    lib.ig_eri_3c_fill.argtypes = [_vp, _vp, ctypes.c_int, i64, i64, _dp, sz, _vp]
    lib.ig_packed_eri_n_ao_pairs.argtypes = [_vp]
    lib.ig_packed_eri_n_ao_pairs.restype = sz
    lib.ig_packed_eri_output_len.argtypes = [_vp]
    lib.ig_packed_eri_output_len.restype = sz
    lib.ig_packed_eri_ao_pairs.argtypes = [_vp, ctypes.POINTER(ctypes.c_int32), sz]
    lib.ig_packed_eri_rows.argtypes = [_vp, i64, i64, ctypes.POINTER(sz), ctypes.POINTER(sz)]
    lib.ig_packed_eri_fill.argtypes = [_vp, ctypes.c_int, i64, i64, ctypes.c_double, _dp, sz, _vp]
    lib.ig_packed_eri_build.argtypes = [_vp, ctypes.c_int, ctypes.c_double, _dp, sz, _vp]
    lib.ig_batch_create.argtypes = [ctypes.POINTER(_vp), sz, ctypes.POINTER(_vp)]
    lib.ig_batch_destroy.argtypes = [_vp]
    lib.ig_batch_destroy.restype = None
    lib.ig_batch_size.argtypes = [_vp]
    lib.ig_batch_size.restype = sz
    lib.ig_batch_set_mixed_precision.argtypes = [_vp, ctypes.c_double]
    lib.ig_batch_jk_device.argtypes = [_vp, ctypes.POINTER(_vp), ctypes.c_double, ctypes.c_int,
                                       ctypes.POINTER(_vp), ctypes.POINTER(_vp), _vp]
    lib.ig_batch_last_stats.argtypes = [_vp, ctypes.POINTER(i64)]
    lib.ig_batch_check.argtypes = [_vp, ctypes.c_double, ctypes.c_int, ctypes.c_char_p, sz]
    lib.ig_batch_autotune.argtypes = [_vp, sz, _vp, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                      _vp]
    lib.ig_batch_schedule_write.argtypes = [_vp, ctypes.POINTER(_ScheduleEntry), sz]
    lib.ig_grad_create.argtypes = [_vp, ctypes.c_int, ctypes.POINTER(_vp)]
    lib.ig_grad_destroy.argtypes = [_vp]
    lib.ig_grad_natom.argtypes = [_vp]
    lib.ig_grad_natom.restype = sz
    lib.ig_grad_nao.argtypes = [_vp]
    lib.ig_grad_nao.restype = sz
    lib.ig_grad_density_create.argtypes = [_vp, sz, _dp, _dpp, _dpp, sz, _dp, _dpp, _dpp,
                                           ctypes.POINTER(_vp)]
    lib.ig_grad_density_destroy.argtypes = [_vp]
    lib.ig_grad_accumulate_bra_pair.argtypes = [_vp, _vp, ctypes.c_int32, ctypes.c_int32,
                                                ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                                _dp, sz, _vp]
    lib.ig_grad_contract.argtypes = [_vp, _vp, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                     _dp, sz, _vp]
    lib.ig_grad_last_stats.argtypes = [_vp, ctypes.POINTER(i64)]
    lib.ig_grad_set_profiling.argtypes = [_vp, ctypes.c_int]
    lib.ig_grad_profile_read.argtypes = [_vp, ctypes.POINTER(_ClassProfile), sz,
                                         ctypes.POINTER(sz)]
    lib.ig_grad_workspace_create.argtypes = [_vp, ctypes.POINTER(_vp)]
    lib.ig_grad_workspace_destroy.argtypes = [_vp]
    lib.ig_grad_workspace_destroy.restype = None
    lib.ig_grad_accumulate_bra_pair_ws.argtypes = [_vp, _vp, _vp, ctypes.c_int32, ctypes.c_int32,
                                                   ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                                   _dp, sz, _vp]
    lib.ig_grad_contract_ws.argtypes = [_vp, _vp, _vp, ctypes.c_int, ctypes.c_double,
                                        ctypes.c_double, _dp, sz, _vp]
    lib.ig_eri_grad_contract_device.argtypes = [_vp, ctypes.c_int, _vp, sz, ctypes.c_int,
                                                ctypes.c_double, _dp, sz, _vp]
    _lib = lib
    return lib


def abi_version():
    return int(library().integral_abi_version())


def _raise(rc):
    if rc == OK:
        return
    msg = library().ig_last_error_message().decode()
    cls = _ERRORS.get(rc)
    if cls is None:
        raise UsageError("%s (status %d)" % (msg, rc))
    raise cls(msg)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _ptr(a):
    return a.ctypes.data_as(_dp)


def _ptrs(arrs):
    return (_dp * len(arrs))(*[_ptr(a) for a in arrs]) if arrs else None


def _stream_ptr(stream):
    if stream is None:
        return None
    return ctypes.c_void_p(int(getattr(stream, "cuda_stream", stream)))


def _check_device_jk(nao, dens, J=None, K=None):
    """Shape checks of the device J/K entry points, done before any library
    call (the kernels trust ndens x nao x nao): dens is a contiguous float64
    CUDA tensor (ndens, nao, nao) with 1 <= ndens <= 64 (direct.rs:141-148),
    J and K (when given) have its shape, dtype and device."""
    def ok(t):
        return (getattr(t, "is_cuda", False) and str(getattr(t, "dtype", "")) == "torch.float64"
                and t.is_contiguous())
    if not ok(dens):
        raise UsageError("densities must be a contiguous float64 CUDA tensor")
    shape = tuple(dens.shape)
    if len(shape) != 3 or shape[1:] != (nao, nao):
        raise UsageError("densities must have shape (ndens, %d, %d), got %s" % (nao, nao, shape))
    if not 1 <= shape[0] <= 64:
        raise UsageError("between 1 and 64 densities, got %d" % shape[0])
    for name, t in (("J", J), ("K", K)):
        if t is None:
            continue
        if not ok(t) or tuple(t.shape) != shape or t.device != dens.device:
            raise UsageError("%s must be a contiguous float64 CUDA tensor of shape %s on %s"
                             % (name, shape, dens.device))


def boys(m_max, T):
    """F_0..F_m_max at the points T, evaluated on the GPU (n x (m_max+1))."""
    T = _f64(np.atleast_1d(T))
    out = np.zeros((T.size, m_max + 1))
    _raise(library().ig_boys_eval(int(m_max), _ptr(T), T.size, _ptr(out)))
    return out


def rys(nroots, T):
    """Rys roots (squared) and weights of the nroots-point rule at the points
    T, evaluated on the GPU (two n x nroots arrays)."""
    T = _f64(np.atleast_1d(T))
    x = np.zeros((T.size, nroots))
    w = np.zeros((T.size, nroots))
    _raise(library().ig_rys_eval(int(nroots), _ptr(T), T.size, _ptr(x), _ptr(w)))
    return x, w


class Shell:
    """A contracted Gaussian shell (shell.rs:49-118).  Validation happens in
    the library when the shell joins a Basis; the same checks run here first
    so construction fails like Shell::with_kind."""

    def __init__(self, l, center, exponents, coefficients, kind=ShellKind.Cartesian):
        exponents = [float(x) for x in exponents]
        coefficients = [float(x) for x in coefficients]
        if not exponents:
            raise EmptyContraction("empty contraction")
        if len(exponents) != len(coefficients):
            raise MismatchedContraction("%d exponents vs %d coefficients"
                                        % (len(exponents), len(coefficients)))
        if l > MAX_L:
            raise AngularMomentumTooHigh("l=%d exceeds max %d" % (l, MAX_L))
        self.l = int(l)
        self.center = tuple(float(c) for c in center)
        self.exponents = exponents
        self.coefficients = coefficients
        self.kind = ShellKind(kind)

    @classmethod
    def new(cls, l, center, exponents, coefficients):
        return cls(l, center, exponents, coefficients, ShellKind.Cartesian)

    @classmethod
    def new_spherical(cls, l, center, exponents, coefficients):
        return cls(l, center, exponents, coefficients, ShellKind.Spherical)

    def n_cart(self):
        return (self.l + 1) * (self.l + 2) // 2

    def n_func(self):
        return 2 * self.l + 1 if self.kind == ShellKind.Spherical else self.n_cart()

    def n_prim(self):
        return len(self.exponents)


class Basis:
    """Basis::new (shell.rs:198-275) backed by an ig_basis handle."""

    def __init__(self, shells):
        lib = library()
        self._shells = list(shells)
        descs = (_ShellDesc * max(len(self._shells), 1))()
        self._keep = []
        for d, s in zip(descs, self._shells):
            e, c = _f64(s.exponents), _f64(s.coefficients)
            self._keep += [e, c]
            d.l, d.kind = s.l, int(s.kind)
            d.center[:] = list(s.center)
            d.n_exponents, d.n_coefficients = len(s.exponents), len(s.coefficients)
            d.exponents, d.coefficients = _ptr(e), _ptr(c)
        h = _vp()
        _raise(lib.ig_basis_create(descs, len(self._shells), ctypes.byref(h)))
        self._h = h
        self._keep = None

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.ig_basis_destroy(self._h)
            self._h = None

    def shells(self):
        return list(self._shells)

    def nao(self):
        return int(library().ig_basis_nao(self._h))

    def nao_cart(self):
        return int(library().ig_basis_nao_cart(self._h))

    def atoms(self):
        seen = []
        for s in self._shells:
            if s.center not in seen:
                seen.append(s.center)
        return seen

    def eri_with(self, engine):
        n = self.nao()
        out = np.zeros(n ** 4)
        _raise(library().ig_eri_dense(self._h, int(engine), _ptr(out), out.size, None))
        return out

    def eri(self):
        return self.eri_with(Engine.Auto)

    def eri_kernel(self, k):
        """Basis::eri_kernel (integrals.rs:773-830): the dense tensor of the
        Coulomb or erf(omega r)/r operator."""
        n = self.nao()
        out = np.zeros(n ** 4)
        _raise(library().ig_eri_kernel(self._h, *k._args(), _ptr(out), out.size, None))
        return out

    def eri_screened_with(self, engine, tau):
        n = self.nao()
        out = np.zeros(n ** 4)
        st = (ctypes.c_int64 * 2)()
        _raise(library().ig_eri_screened(self._h, int(engine), float(tau), _ptr(out), out.size,
                                         st, None))
        return out, {"shell_quartets_total": st[0], "shell_quartets_skipped": st[1],
                     "tau": float(tau)}

    def eri_screened(self, tau):
        return self.eri_screened_with(Engine.Auto, tau)

    def schwarz_bounds_with(self, engine):
        nsh = len(self._shells)
        out = np.zeros(nsh * nsh)
        _raise(library().ig_schwarz_bounds(self._h, int(engine), _ptr(out), out.size, None))
        return out

    def schwarz_bounds(self):
        return self.schwarz_bounds_with(Engine.Auto)

    def direct_contractor(self):
        return DirectContractor(self)

    def packed_eri_builder(self):
        return PackedEriBuilder(self)

    def eri_builder(self):
        return EriBuilder(self)

    # density fitting (df.rs): self as the auxiliary basis for 2c, as the
    # main basis for 3c
    def eri_2c_with(self, engine):
        n = self.nao()
        out = np.zeros(n * n)
        _raise(library().ig_eri_2c(self._h, int(engine), _ptr(out), out.size, None))
        return out

    def eri_2c(self):
        return self.eri_2c_with(Engine.Auto)

    def eri_2c_kernel(self, k):
        """Basis::eri_2c_kernel (df.rs:95-128)."""
        n = self.nao()
        out = np.zeros(n * n)
        _raise(library().ig_eri_2c_kernel(self._h, *k._args(), _ptr(out), out.size, None))
        return out

    def schwarz_aux_bounds_with(self, engine):
        out = np.zeros(len(self._shells))
        _raise(library().ig_schwarz_aux_bounds(self._h, int(engine), _ptr(out), out.size, None))
        return out

    def schwarz_aux_bounds(self):
        return self.schwarz_aux_bounds_with(Engine.Auto)

    def eri_3c_builder(self, aux):
        return Eri3cBuilder(self, aux)

    def eri_3c_block_with(self, engine, aux, ish, jsh, psh, kernel=None):
        """One (ij|P) block, row-major [n_i, n_j, n_p] (df.rs:148-199)."""
        b = Eri3cBuilder(self, aux, engine)
        b._kernel = kernel or EriKernel.coulomb()
        hi, lo = (ish, jsh) if ish >= jsh else (jsh, ish)
        out = np.zeros(b.output_len())
        b.fill_range(hi * (hi + 1) // 2 + lo, hi * (hi + 1) // 2 + lo + 1, out)
        n, naux = self.nao(), aux.nao()
        off = np.concatenate([[0], np.cumsum([s.n_func() for s in self._shells])])
        aoff = np.concatenate([[0], np.cumsum([s.n_func() for s in aux._shells])])
        t = out.reshape(n, n, naux)
        return t[off[ish]:off[ish + 1], off[jsh]:off[jsh + 1], aoff[psh]:aoff[psh + 1]].ravel()

    def eri_3c_block(self, aux, ish, jsh, psh):
        return self.eri_3c_block_with(Engine.Auto, aux, ish, jsh, psh)

    def eri_3c_block_kernel(self, aux, ish, jsh, psh, k):
        """Basis::eri_3c_block_kernel (df.rs:190-240)."""
        return self.eri_3c_block_with(Engine.Auto, aux, ish, jsh, psh, kernel=k)

    def eri_grad_contract_with(self, engine, gamma, kernel=None, stream=None):
        """Basis::eri_grad_contract_with (grad.rs:426-466): per atom
        F_c = sum over all ordered quartets of gamma[mu,nu,lam,sig] *
        d(mu nu|lam sig)/dR_c, returned as (natom, 3) in atoms() order.
        gamma is the dense nao^4 row-major two-particle density, a float64
        CUDA tensor (used in place) or a host array (copied to the device).
        `kernel` as in eri_grad_contract_kernel (grad.rs:493-540)."""
        import torch
        if isinstance(gamma, torch.Tensor):
            g = gamma
            if not (g.is_cuda and g.dtype == torch.float64 and g.is_contiguous()):
                raise UsageError("gamma must be a contiguous float64 CUDA tensor")
        else:
            g = torch.from_numpy(np.ascontiguousarray(gamma, dtype=np.float64).ravel()).to("cuda")
        kernel = kernel or EriKernel.coulomb()
        out = np.zeros((len(self.atoms()), 3))
        _raise(library().ig_eri_grad_contract_device(
            self._h, int(engine), _vp(g.data_ptr()), g.numel(), *kernel._args(), _ptr(out),
            out.size, _stream_ptr(stream)))
        return out

    def eri_grad_contract(self, gamma):
        """Basis::eri_grad_contract (grad.rs:417-419)."""
        return self.eri_grad_contract_with(Engine.Auto, gamma)

    def eri_grad_contract_kernel(self, gamma, k):
        """Basis::eri_grad_contract_kernel (grad.rs:493-540): Coulomb or erf."""
        return self.eri_grad_contract_with(Engine.Auto, gamma, kernel=k)

    def direct_grad_contractor(self):
        return DirectGradContractor(self)


class BraPairFill:
    """eri_builder.rs BraPairFill: one canonical bra pair's task over the
    caller's nao^4 buffer (the GPU fill writes only the rows it owns)."""

    def __init__(self, bra, out):
        self._bra = bra
        self.out = out

    def bra(self):
        return self._bra


class EriBuilder:
    """EriBuilder (eri_builder.rs:101-454): the dense nao^4 tensor filled by
    canonical bra-pair tasks that own disjoint output rows."""

    def __init__(self, basis, engine=Engine.Auto):
        self._basis = basis
        self._engine = int(engine)
        self._kernel = EriKernel.coulomb()
        nsh = len(basis._shells)
        self._pairs = [(i, j) for i in range(nsh) for j in range(i + 1)]

    @classmethod
    def with_engine(cls, basis, engine):
        return cls(basis, engine)

    def kernel(self, k):
        """EriBuilder::kernel (eri_builder.rs:280-305): the operator of every
        later fill/build; erf needs a finite omega > 0 (checked here, as the
        reference's panic)."""
        if k.omega is not None and not (math.isfinite(k.omega) and k.omega > 0.0):
            raise UsageError("EriKernel::Erf requires a finite omega > 0")
        self._kernel = k
        return self

    def bra_pairs(self):
        return list(self._pairs)

    def output_len(self):
        return self._basis.nao() ** 4

    def _check(self, out):
        if not (isinstance(out, np.ndarray) and out.dtype == np.float64 and out.flags.c_contiguous
                and out.size == self.output_len()):
            raise UsageError("ERI output buffer must hold %d elements" % self.output_len())

    def partition(self, out):
        """One task per canonical bra pair (eri_builder.rs:337-399)."""
        self._check(out)
        return [BraPairFill(k, out) for k in range(len(self._pairs))]

    def fill(self, task, stream=None):
        self.fill_range(task.bra(), task.bra() + 1, task.out, stream)

    def fill_range(self, bra_begin, bra_end, out, stream=None):
        """The rows of the canonical bra pairs [bra_begin, bra_end) at once."""
        self._check(out)
        _raise(library().ig_eri_builder_fill_kernel(self._basis._h, self._engine,
                                                    *self._kernel._args(), int(bra_begin),
                                                    int(bra_end), _ptr(out), out.size,
                                                    _stream_ptr(stream)))

    def build(self):
        if self._kernel.omega is None:
            return self._basis.eri_with(self._engine)
        return self._basis.eri_kernel(self._kernel)


class Bra3cFill:
    """df.rs:531-548: the (mu, nu) rows one canonical bra pair owns."""

    def __init__(self, bra, out):
        self._bra = bra
        self.out = out

    def bra(self):
        return self._bra


class Eri3cBuilder:
    """Eri3cBuilder (df.rs:325-530): the dense [nao, nao, naux] 3-center
    tensor, P fastest, filled by canonical main-basis bra-pair tasks."""

    def __init__(self, main, aux, engine=Engine.Auto):
        self._main, self._aux, self._engine = main, aux, int(engine)
        self._kernel = EriKernel.coulomb()
        nsh = len(main._shells)
        self._pairs = [(i, j) for i in range(nsh) for j in range(i + 1)]

    @classmethod
    def with_engine(cls, main, aux, engine):
        return cls(main, aux, engine)

    def bra_pairs(self):
        return list(self._pairs)

    def output_len(self):
        return self._main.nao() ** 2 * self._aux.nao()

    def _check(self, out):
        if not (isinstance(out, np.ndarray) and out.dtype == np.float64 and out.flags.c_contiguous
                and out.size == self.output_len()):
            raise UsageError("3c output buffer must be nao^2*naux = %d elements"
                             % self.output_len())

    def partition(self, out):
        self._check(out)
        return [Bra3cFill(p, out) for p in self._pairs]

    def fill(self, task, stream=None):
        i, j = task.bra()
        k = i * (i + 1) // 2 + j
        self.fill_range(k, k + 1, task.out, stream)

    def fill_range(self, bra_begin, bra_end, out, stream=None):
        self._check(out)
        _raise(library().ig_eri_3c_fill_kernel(self._main._h, self._aux._h, self._engine,
                                               *self._kernel._args(), int(bra_begin),
                                               int(bra_end), _ptr(out), out.size,
                                               _stream_ptr(stream)))

    def build(self):
        out = np.zeros(self.output_len())
        self.fill_range(0, -1, out)
        return out


class PackedEriWorkspace:
    """packed_eri.rs:34-47.  The device plan owns all scratch; kept for API parity."""


class PackedBraFill:
    """packed_eri.rs:51-63: the packed rows owned by one canonical bra pair."""

    def __init__(self, bra, rows):
        self._bra = bra
        self.rows = rows

    def bra(self):
        return self._bra


class PackedEriBuilder:
    """PackedEriBuilder (packed_eri.rs:67-240): the nao^4/8 symmetry-unique
    ERIs over shell-pair-blocked AO pairs, (t|u) at t(t+1)/2 + u."""

    def __init__(self, basis, engine=Engine.Auto):
        self._basis = basis
        self._engine = int(engine)
        nsh = len(basis._shells)
        self._pairs = [(i, j) for i in range(nsh) for j in range(i + 1)]

    @classmethod
    def with_engine(cls, basis, engine):
        return cls(basis, engine)

    def bra_pairs(self):
        return list(self._pairs)

    def ao_pairs(self):
        lib = library()
        n = int(lib.ig_packed_eri_n_ao_pairs(self._basis._h))
        buf = (ctypes.c_int32 * max(2 * n, 1))()
        _raise(lib.ig_packed_eri_ao_pairs(self._basis._h, buf, n))
        return [(buf[2 * k], buf[2 * k + 1]) for k in range(n)]

    def output_len(self):
        return int(library().ig_packed_eri_output_len(self._basis._h))

    def _rows(self, b0, b1):
        off, ln = ctypes.c_size_t(), ctypes.c_size_t()
        _raise(library().ig_packed_eri_rows(self._basis._h, int(b0), int(b1), ctypes.byref(off),
                                            ctypes.byref(ln)))
        return off.value, ln.value

    def partition(self, out):
        """One task per bra pair over disjoint slices of `out` (a zeroed
        float64 array of output_len()); packed_eri.rs:146-171."""
        if not (isinstance(out, np.ndarray) and out.dtype == np.float64 and out.flags.c_contiguous
                and out.size == self.output_len()):
            raise UsageError("packed ERI output buffer must hold %d elements" % self.output_len())
        tasks = []
        for bra in range(len(self._pairs)):
            off, ln = self._rows(bra, bra + 1)
            tasks.append(PackedBraFill(bra, out[off:off + ln]))
        return tasks

    def fill(self, workspace, task, threshold=0.0, stream=None):
        """Evaluate one bra pair's quartets into its rows (packed_eri.rs:173-220)."""
        rows = task.rows
        _raise(library().ig_packed_eri_fill(self._basis._h, self._engine, task.bra(),
                                            task.bra() + 1, float(threshold), _ptr(rows),
                                            rows.size, _stream_ptr(stream)))

    def fill_range(self, bra_begin, bra_end, rows, threshold=0.0, stream=None):
        """fill() for the contiguous bra pairs [bra_begin, bra_end) at once."""
        _raise(library().ig_packed_eri_fill(self._basis._h, self._engine, int(bra_begin),
                                            int(bra_end), float(threshold), _ptr(rows), rows.size,
                                            _stream_ptr(stream)))

    def build(self, threshold=0.0, stream=None):
        out = np.zeros(self.output_len())
        _raise(library().ig_packed_eri_build(self._basis._h, self._engine, float(threshold),
                                             _ptr(out), out.size, _stream_ptr(stream)))
        return out


class _Workspace:
    """Per-worker device scratch bound to one contractor on first use (the
    reference's workspaces grow on demand and serve any contractor; here a
    workspace used with another contractor rebinds to it)."""

    _create = _destroy = None

    def __init__(self):
        self._h = None
        self._owner = None

    def _handle(self, contractor):
        if self._owner is not contractor:
            self._release()
            h = _vp()
            _raise(getattr(library(), self._create)(contractor._h, ctypes.byref(h)))
            self._h, self._owner = h, contractor
        return self._h

    def _release(self):
        if self._h is not None and _lib is not None:
            getattr(_lib, self._destroy)(self._h)
        self._h = self._owner = None

    def __del__(self):
        self._release()


class DirectWorkspace(_Workspace):
    """DirectWorkspace (direct.rs:19-33): a worker's own device scratch, so
    workers sharing one DirectContractor run concurrently."""

    _create, _destroy = "ig_direct_workspace_create", "ig_direct_workspace_destroy"

    def last_stats(self):
        if self._h is None:
            return 0, 0
        st = (ctypes.c_int64 * 2)()
        _raise(library().ig_direct_workspace_last_stats(self._h, st))
        return int(st[0]), int(st[1])


class DirectBuffers:
    """direct.rs:14-17: caller-owned J (and optionally K), accumulated."""

    def __init__(self, coulomb, exchange=None):
        self.coulomb = coulomb
        self.exchange = exchange


class DirectContractor:
    """DirectContractor (direct.rs:39-218) over an ig_direct plan."""

    def __init__(self, basis, engine=Engine.Auto):
        h = _vp()
        _raise(library().ig_direct_create(basis._h, int(engine), ctypes.byref(h)))
        self._h = h
        self._nsh = len(basis._shells)

    @classmethod
    def with_engine(cls, basis, engine):
        return cls(basis, engine)

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.ig_direct_destroy(self._h)
            self._h = None

    def nao(self):
        return int(library().ig_direct_nao(self._h))

    def bra_pairs(self):
        n = int(library().ig_direct_n_bra_pairs(self._h))
        buf = (ctypes.c_int32 * (2 * n))()
        _raise(library().ig_direct_bra_pairs(self._h, buf, n))
        return [(buf[2 * k], buf[2 * k + 1]) for k in range(n)]

    def density_shell_bounds(self, densities):
        ds = [_f64(d).ravel() for d in densities]
        out = np.zeros(len(ds) * self._nsh * self._nsh)
        _raise(library().ig_direct_density_shell_bounds(self._h, _ptrs(ds), len(ds), _ptr(out),
                                                        out.size))
        return [out[k * self._nsh ** 2:(k + 1) * self._nsh ** 2] for k in range(len(ds))]

    @staticmethod
    def _check_out(arrs, n2):
        for a in arrs:
            if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and a.flags.c_contiguous
                    and a.size == n2):
                raise UsageError("J/K buffers must be C-contiguous float64 arrays of nao*nao")

    def accumulate_bra_pair(self, workspace, bra, densities, density_bounds, threshold,
                            want_exchange, outputs, stream=None):
        ds = [_f64(d).ravel() for d in densities]
        if len(outputs) != len(ds) or len(density_bounds) != len(ds):
            raise UsageError("densities, bounds and outputs differ in length")
        n2 = self.nao() ** 2
        js = [o.coulomb for o in outputs]
        ks = [o.exchange for o in outputs] if want_exchange else []
        self._check_out(js + ks, n2)
        bounds = _f64(np.concatenate([np.ravel(b) for b in density_bounds])) if ds else _f64([0.0])
        ws = workspace._handle(self) if isinstance(workspace, DirectWorkspace) else None
        _raise(library().ig_direct_accumulate_bra_pair_ws(
            self._h, ws, int(bra[0]), int(bra[1]), _ptrs(ds), _ptr(bounds), len(ds),
            float(threshold), 1 if want_exchange else 0, _ptrs(js),
            _ptrs(ks) if want_exchange else None, _stream_ptr(stream)))

    def jk(self, densities, threshold=0.0, want_exchange=True, bra_range=(0, -1), J=None, K=None,
           stream=None, workspace=None):
        """Bulk J/K over canonical bra pairs [bra_range) (all by default);
        accumulates into J/K when given, else returns fresh matrices."""
        ds = [_f64(d).ravel() for d in densities]
        n = self.nao()
        J = [np.zeros(n * n) for _ in ds] if J is None else J
        K = ([np.zeros(n * n) for _ in ds] if K is None else K) if want_exchange else None
        self._check_out(J + (K or []), n * n)
        ws = workspace._handle(self) if workspace is not None else None
        _raise(library().ig_direct_jk_ws(self._h, ws, _ptrs(ds), len(ds), float(threshold),
                                         1 if want_exchange else 0, _ptrs(J),
                                         _ptrs(K) if want_exchange else None, int(bra_range[0]),
                                         int(bra_range[1]), _stream_ptr(stream)))
        return J, K

    def jk_device(self, dens, J, K=None, threshold=0.0, stream=None, workspace=None,
                  symmetric=None):
        """Device-resident J/K: dens, J, K are torch.float64 CUDA tensors of
        shape (ndens, nao, nao); J/K are accumulated.  With symmetric=None
        the densities' symmetry is tested on the device (one 4-byte read
        back); True/False declares it and the call never synchronises.
        `workspace` (a DirectWorkspace) gives the call its own scratch."""
        _check_device_jk(self.nao(), dens, J, K)
        if symmetric is None and workspace is None:
            _raise(library().ig_direct_jk_device(
                self._h, _vp(dens.data_ptr()), dens.shape[0], float(threshold),
                1 if K is not None else 0, _vp(J.data_ptr()),
                _vp(K.data_ptr()) if K is not None else None, _stream_ptr(stream)))
            return
        ws = workspace._handle(self) if workspace is not None else None
        _raise(library().ig_direct_jk_device_ws(
            self._h, ws, _vp(dens.data_ptr()), dens.shape[0], float(threshold),
            1 if K is not None else 0, _vp(J.data_ptr()),
            _vp(K.data_ptr()) if K is not None else None, 0 if symmetric is False else 1,
            _stream_ptr(stream)))

    def jk_device_noalloc(self, dens, J, K=None, threshold=0.0, symmetric=True, stream=None):
        """jk_device with no device allocation: the call's scratch comes from
        the library's fixed 64 MiB device array (shared by every call on the
        GPU: overlapping calls race).  No statistics are recorded."""
        _check_device_jk(self.nao(), dens, J, K)
        _raise(library().ig_direct_jk_device_noalloc(
            self._h, _vp(dens.data_ptr()), dens.shape[0], float(threshold),
            1 if K is not None else 0, _vp(J.data_ptr()),
            _vp(K.data_ptr()) if K is not None else None, 1 if symmetric else 0,
            _stream_ptr(stream)))

    def last_stats(self):
        st = (ctypes.c_int64 * 2)()
        _raise(library().ig_direct_last_stats(self._h, st))
        return int(st[0]), int(st[1])

    def set_profiling(self, enable):
        _raise(library().ig_direct_set_profiling(self._h, 1 if enable else 0))

    # J/K kernel variants of a compiled quartet class
    TEAM, REGISTER, ROW_ACCUMULATE, OS_HGP, TEAM_HRR = 0, 1, 2, 3, 4

    def autotune(self, dens, threshold=0.0, want_exchange=True, reps=2, stream=None):
        """Time every J/K kernel variant of every class pair on the device
        densities `dens` (float64 CUDA tensor, ndens x nao x nao) and keep the
        faster per class pair; returns the resulting schedule."""
        _check_device_jk(self.nao(), dens)
        _raise(library().ig_direct_autotune(self._h, _vp(dens.data_ptr()), dens.shape[0],
                                            float(threshold), 1 if want_exchange else 0,
                                            int(reps), _stream_ptr(stream)))
        return self.schedule()

    def schedule(self):
        """[{"cls": (la, lb, lc, ld), "variant": v, "tuned": bool, "ms": t}] for
        every compiled class pair of this basis."""
        lib = library()
        n = ctypes.c_size_t()
        _raise(lib.ig_direct_schedule_read(self._h, None, 0, ctypes.byref(n)))
        buf = (_ScheduleEntry * max(n.value, 1))()
        _raise(lib.ig_direct_schedule_read(self._h, buf, n.value, ctypes.byref(n)))
        return [{"cls": (e.la, e.lb, e.lc, e.ld), "variant": e.variant, "tuned": bool(e.tuned),
                 "ms": e.ms} for e in buf[:n.value]]

    def set_schedule(self, entries):
        """Load entries as returned by schedule() (variant -1 = default)."""
        buf = (_ScheduleEntry * max(len(entries), 1))()
        for b, e in zip(buf, entries):
            b.la, b.lb, b.lc, b.ld = e["cls"]
            b.variant = int(e["variant"])
            b.ms = float(e.get("ms", 0.0))
        _raise(library().ig_direct_schedule_write(self._h, buf, len(entries)))

    def profile(self):
        """Per-class records of the last call: dicts with la..ld, quartets,
        prim_quartets and ms (event time when profiling is on)."""
        cap = 1024
        buf = (_ClassProfile * cap)()
        n = ctypes.c_size_t()
        _raise(library().ig_direct_profile_read(self._h, buf, cap, ctypes.byref(n)))
        return [{"cls": (r.la, r.lb, r.lc, r.ld), "fp32": bool(r.fp32), "quartets": r.quartets,
                 "prim_quartets": r.prim_quartets, "ms": r.ms} for r in buf[:min(n.value, cap)]]

    def set_mixed_precision(self, fp32_below):
        """Quartets with Q_ij*Q_kl < fp32_below run with FP32 tables and block
        in the whole-list J/K calls (0 = FP64 only)."""
        _raise(library().ig_direct_set_mixed_precision(self._h, float(fp32_below)))


class BatchContractor:
    """J/K of many molecules in one HIP graph (no counterpart in the
    reference, whose DirectContractor serves one basis, direct.rs:39-131):
    one plan per basis, one persistent launch per quartet class over the
    concatenated quartet lists of every molecule, the sequence captured once
    and replayed while threshold, exchange flag and buffers stay the same."""

    def __init__(self, bases):
        bases = list(bases)
        if not bases:
            raise UsageError("a batch needs at least one basis")
        hs = (_vp * len(bases))(*[b._h for b in bases])
        h = _vp()
        _raise(library().ig_batch_create(hs, len(bases), ctypes.byref(h)))
        self._h = h
        self._bases = bases  # the plans copy them; kept for nao()
        self._nao = [b.nao() for b in bases]

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.ig_batch_destroy(self._h)
            self._h = None

    def __len__(self):
        return int(library().ig_batch_size(self._h))

    def nao(self, m):
        return self._nao[m]

    def set_mixed_precision(self, fp32_below):
        """As DirectContractor.set_mixed_precision, for every molecule."""
        _raise(library().ig_batch_set_mixed_precision(self._h, float(fp32_below)))

    def jk_device(self, dens, J, K=None, threshold=0.0, stream=None):
        """dens, J, K: lists (one per molecule) of contiguous float64 CUDA
        tensors of nao_m x nao_m; the J and K of the symmetrised densities
        are added to J/K.  Asynchronous on `stream`."""
        n = len(self._nao)
        lists = [dens, J] + ([K] if K is not None else [])
        for lst in lists:
            if len(lst) != n:
                raise UsageError("one density and one output per molecule")
            for m, t in enumerate(lst):
                if not (t.is_cuda and t.element_size() == 8 and t.dtype.is_floating_point
                        and t.is_contiguous() and t.numel() == self._nao[m] ** 2):
                    raise UsageError("device buffers must be contiguous float64 CUDA tensors "
                                     "of nao*nao for their molecule")

        def arr(lst):
            return (_vp * n)(*[t.data_ptr() for t in lst])

        _raise(library().ig_batch_jk_device(self._h, arr(dens), float(threshold),
                                            1 if K is not None else 0, arr(J),
                                            arr(K) if K is not None else None,
                                            _stream_ptr(stream)))

    def last_stats(self):
        st = (ctypes.c_int64 * 2)()
        _raise(library().ig_batch_last_stats(self._h, st))
        return int(st[0]), int(st[1])

    def check(self, threshold=0.0, want_exchange=True):
        """Host walk of the batched launches' indexing for this threshold
        (ig_batch_check): runs no kernel; returns "" when every access is in
        range, else the first offending access."""
        buf = ctypes.create_string_buffer(512)
        rc = library().ig_batch_check(self._h, float(threshold), 1 if want_exchange else 0,
                                      buf, len(buf))
        if rc == 0:
            return ""
        msg = buf.value.decode()
        if msg:
            return msg
        _raise(rc)
        return "check failed"

    def set_schedule(self, entries):
        """Load entries as returned by DirectContractor.schedule() into every
        molecule's plan."""
        buf = (_ScheduleEntry * max(len(entries), 1))()
        for b, e in zip(buf, entries):
            b.la, b.lb, b.lc, b.ld = e["cls"]
            b.variant = int(e["variant"])
            b.ms = float(e.get("ms", 0.0))
        _raise(library().ig_batch_schedule_write(self._h, buf, len(entries)))

    def autotune(self, dens, threshold=0.0, want_exchange=True, reps=1, molecule=None,
                 stream=None):
        """Tune the class-to-kernel schedule on one molecule (default: the
        largest) with its density `dens[molecule]` (a list of CUDA tensors as
        for jk_device) and use it for the whole batch."""
        m = int(np.argmax(self._nao)) if molecule is None else int(molecule)
        if not 0 <= m < len(self._nao):
            raise UsageError("molecule out of range")
        t = dens[m]
        if not (t.is_cuda and t.element_size() == 8 and t.is_contiguous()
                and t.numel() == self._nao[m] ** 2):
            raise UsageError("density must be a contiguous float64 CUDA tensor of nao*nao")
        _raise(library().ig_batch_autotune(self._h, m, _vp(t.data_ptr()), float(threshold),
                                           1 if want_exchange else 0, int(reps),
                                           _stream_ptr(stream)))
        return m


class GradDensityTerm:
    """direct_grad.rs:22-26: factor * left (x) right."""

    def __init__(self, factor, left, right):
        self.factor = float(factor)
        self.left = left
        self.right = right


class DirectGradWorkspace(_Workspace):
    """DirectGradWorkspace (direct_grad.rs:51-67): a worker's own device
    scratch over a shared DirectGradContractor."""

    _create, _destroy = "ig_grad_workspace_create", "ig_grad_workspace_destroy"


class GradDensity:
    def __init__(self, owner, handle, keep):
        self._owner = owner
        self._h = handle
        self._keep = keep

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.ig_grad_density_destroy(self._h)
            self._h = None


class DirectGradContractor:
    """DirectGradContractor (direct_grad.rs:144-392) over an ig_grad plan."""

    def __init__(self, basis, engine=Engine.Auto):
        h = _vp()
        _raise(library().ig_grad_create(basis._h, int(engine), ctypes.byref(h)))
        self._h = h
        self._nsh = len(basis._shells)

    @classmethod
    def with_engine(cls, basis, engine):
        return cls(basis, engine)

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.ig_grad_destroy(self._h)
            self._h = None

    def natom(self):
        return int(library().ig_grad_natom(self._h))

    def nao(self):
        return int(library().ig_grad_nao(self._h))

    def bra_pairs(self):
        return [(i, j) for i in range(self._nsh) for j in range(i + 1)]

    def density(self, coulomb, exchange):
        def unpack(terms):
            f = _f64([t.factor for t in terms] or [0.0])
            left = [_f64(t.left).ravel() for t in terms]
            right = [_f64(t.right).ravel() for t in terms]
            return f, left, right
        cf, cl, cr = unpack(coulomb)
        xf, xl, xr = unpack(exchange)
        n2 = self.nao() ** 2
        for m in cl + cr + xl + xr:
            if m.size != n2:
                raise UsageError("density matrices must be nao x nao")
        h = _vp()
        _raise(library().ig_grad_density_create(self._h, len(coulomb), _ptr(cf), _ptrs(cl),
                                                _ptrs(cr), len(exchange), _ptr(xf), _ptrs(xl),
                                                _ptrs(xr), ctypes.byref(h)))
        return GradDensity(self, h, None)

    def accumulate_bra_pair(self, workspace, bra, density, kernel, threshold, out, stream=None):
        out = np.asarray(out)
        if out.dtype != np.float64 or not out.flags.c_contiguous:
            raise UsageError("out must be a C-contiguous float64 array of natom x 3")
        kernel = kernel or EriKernel.coulomb()
        ws = workspace._handle(self) if isinstance(workspace, DirectGradWorkspace) else None
        _raise(library().ig_grad_accumulate_bra_pair_ws(self._h, ws, density._h, int(bra[0]),
                                                        int(bra[1]), *kernel._args(),
                                                        float(threshold), _ptr(out), out.size,
                                                        _stream_ptr(stream)))

    def contract(self, density, kernel=None, threshold=0.0, stream=None, workspace=None):
        kernel = kernel or EriKernel.coulomb()
        out = np.zeros((self.natom(), 3))
        ws = workspace._handle(self) if workspace is not None else None
        _raise(library().ig_grad_contract_ws(self._h, ws, density._h, *kernel._args(),
                                             float(threshold), _ptr(out), out.size,
                                             _stream_ptr(stream)))
        return out

    def last_stats(self):
        st = (ctypes.c_int64 * 2)()
        _raise(library().ig_grad_last_stats(self._h, st))
        return int(st[0]), int(st[1])

    def set_profiling(self, enable):
        _raise(library().ig_grad_set_profiling(self._h, 1 if enable else 0))

    def profile(self):
        """Per-class records of the last contraction (as
        DirectContractor.profile)."""
        cap = 1024
        buf = (_ClassProfile * cap)()
        n = ctypes.c_size_t()
        _raise(library().ig_grad_profile_read(self._h, buf, cap, ctypes.byref(n)))
        return [{"cls": (r.la, r.lb, r.lc, r.ld), "quartets": r.quartets,
                 "prim_quartets": r.prim_quartets, "ms": r.ms} for r in buf[:min(n.value, cap)]]
