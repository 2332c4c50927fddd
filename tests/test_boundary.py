"""CPU suite: the C-ABI library builds, loads and exports every declared
symbol; host-side validation mirrors the reference's errors; helpers that the
benchmark relies on are deterministic.  No compute call needs a GPU here."""
import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import integral_amd as ig
from integral_amd import flops, synthetic

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "integral_hip.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\**\s*(\w+)\s*\(", src, flags=re.M)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(ig.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert ig.abi_version() == 1


def test_no_cpu_fallback_without_gpu():
    lib = ig.library()
    if lib.ig_device_count() > 0:
        pytest.skip("a GPU is visible")
    s = synthetic.to_basis([(0, (0.0, 0.0, 0.0), [1.0], [1.0], False)])
    with pytest.raises(ig.DeviceError):
        s.eri()
    with pytest.raises(ig.DeviceError):
        s.direct_contractor()


def desc(l=0, kind=0, ne=1, nc=1):
    e = (ctypes.c_double * 2)(1.0, 0.5)
    c = (ctypes.c_double * 2)(1.0, 0.5)
    d = ig._ShellDesc()
    d.l, d.kind = l, kind
    d.center[:] = [0.0, 0.0, 0.0]
    d.n_exponents, d.n_coefficients = ne, nc
    d.exponents = ctypes.cast(e, ig._dp)
    d.coefficients = ctypes.cast(c, ig._dp)
    return d, (e, c)


@pytest.mark.parametrize("kw,code", [
    ({"ne": 0, "nc": 0}, 1),          # IntegralError::EmptyContraction
    ({"ne": 2, "nc": 1}, 2),          # IntegralError::MismatchedContraction
    ({"l": 7}, 3),                    # IntegralError::AngularMomentumTooHigh
    ({"kind": 5}, 16),                # bad ShellKind
])
def test_basis_create_status_codes(kw, code):
    d, keep = desc(**kw)
    arr = (ig._ShellDesc * 1)(d)
    h = ctypes.c_void_p()
    rc = ig.library().ig_basis_create(arr, 1, ctypes.byref(h))
    assert rc == code
    assert ig.library().ig_last_error_message()


def test_python_shell_validation_mirrors_reference():
    with pytest.raises(ig.EmptyContraction):
        ig.Shell(0, (0, 0, 0), [], [])
    with pytest.raises(ig.MismatchedContraction):
        ig.Shell(0, (0, 0, 0), [1.0, 2.0], [1.0])
    with pytest.raises(ig.AngularMomentumTooHigh):
        ig.Shell(7, (0, 0, 0), [1.0], [1.0])


def test_basis_dimensions_and_atoms():
    b = synthetic.to_basis([(2, (0.0, 0.0, 0.0), [1.0], [1.0], True),
                            (1, (0.0, 0.0, 0.0), [1.0], [1.0], False),
                            (3, (1.0, 0.0, 0.0), [1.0], [1.0], True)])
    assert b.nao() == 5 + 3 + 7 and b.nao_cart() == 6 + 3 + 10
    assert ig.library().ig_basis_natom(b._h) == 2


def test_gradient_rejects_l6_before_any_device_work():
    b = synthetic.to_basis([(6, (0.0, 0.0, 0.0), [1.0], [1.0], True)])
    with pytest.raises(ig.AngularMomentumTooHighForGradient):
        b.direct_grad_contractor()


def test_dense_gamma_gradient_checks_before_any_device_work():
    """ig_eri_grad_contract_device validates l, gamma length and omega before
    it touches the (here dummy, never dereferenced) device pointer."""
    lib = ig.library()
    fake = ctypes.c_void_p(256)
    out = np.zeros(3)
    b6 = synthetic.to_basis([(6, (0.0, 0.0, 0.0), [1.0], [1.0], True)])
    assert lib.ig_eri_grad_contract_device(b6._h, 0, fake, 13 ** 4, 0, 0.0, ig._ptr(out), 3,
                                           None) == ig.E_LGRAD
    b = synthetic.to_basis([(1, (0.0, 0.0, 0.0), [1.0], [1.0], False)])
    assert lib.ig_eri_grad_contract_device(b._h, 0, fake, 80, 0, 0.0, ig._ptr(out), 3,
                                           None) == ig.E_GAMMA
    assert lib.ig_eri_grad_contract_device(b._h, 0, fake, 81, 1, float("nan"), ig._ptr(out), 3,
                                           None) == 15
    assert lib.ig_eri_grad_contract_device(b._h, 0, None, 81, 0, 0.0, ig._ptr(out), 3,
                                           None) == 16


def test_invalid_engine_and_null_arguments():
    b = synthetic.to_basis([(0, (0.0, 0.0, 0.0), [1.0], [1.0], False)])
    h = ctypes.c_void_p()
    assert ig.library().ig_direct_create(b._h, 9, ctypes.byref(h)) == 16
    assert ig.library().ig_direct_create(None, 0, ctypes.byref(h)) == 16


def test_synthetic_workloads_are_deterministic():
    s1, _ = synthetic.dense_eri_config()
    s1b, _ = synthetic.dense_eri_config()
    assert s1 == s1b
    assert sum((l + 1) * (l + 2) // 2 for (l, *_rest) in s1) == 50
    s2, dens, tau = synthetic.direct_jk_config()
    assert sum(2 * s[0] + 1 for s in s2) == 520 and len(dens) == 2 and tau == 1e-12
    assert np.allclose(dens[0], dens[0].T)
    s4, D = synthetic.gradient_config()
    assert sum(2 * s[0] + 1 for s in s4) == 390
    rng = synthetic.SplitMix64(1)
    assert [rng.next_u64() for _ in range(2)] == [0x910A2DEC89025CC1, 0xBEEB8DA1658EEC67]


def test_flop_model_reference_values():
    # per-primitive-quartet Rys flops quoted for the reference engines
    assert [flops.rys_prim(*ls) for ls in [(0, 0, 0, 0), (1, 0, 1, 0), (1, 1, 1, 1),
                                           (2, 2, 2, 2), (3, 3, 3, 3)]] == \
        [111, 339, 1819, 30975, 299877]
    assert flops.rys_prim(6, 6, 6, 6) == 32270679


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr


class _FakeTensor:
    """Just the metadata the device-argument checks read (no GPU needed)."""

    def __init__(self, shape, dtype="torch.float64", cuda=True, contiguous=True, device="cuda:0"):
        self.shape, self.dtype, self.is_cuda, self.device = tuple(shape), dtype, cuda, device
        self._contiguous = contiguous

    def is_contiguous(self):
        return self._contiguous


@pytest.mark.parametrize("dens,J,K,ok", [
    ((2, 5, 5), (2, 5, 5), (2, 5, 5), True),
    ((2, 5, 5), (2, 5, 5), None, True),
    ((5, 5), (5, 5), (5, 5), False),             # a 2-D density would read as 5 densities
    ((2, 25), (2, 25), (2, 25), False),
    ((2, 5, 5), (1, 5, 5), (2, 5, 5), False),    # J smaller than the densities
    ((2, 5, 5), (2, 5, 5), (2, 4, 5), False),
    ((0, 5, 5), (0, 5, 5), None, False),         # direct.rs:141-144: at least one density
    ((65, 5, 5), (65, 5, 5), None, False),       # direct.rs:145-148: at most 64
])
def test_device_jk_shapes_are_checked_before_any_library_call(dens, J, K, ok):
    args = [_FakeTensor(dens), _FakeTensor(J), _FakeTensor(K) if K else None]
    if ok:
        ig._check_device_jk(5, *args)
    else:
        with pytest.raises(ig.UsageError):
            ig._check_device_jk(5, *args)


def test_device_jk_dtype_device_and_layout_are_checked():
    good = (2, 5, 5)
    for bad in (_FakeTensor(good, dtype="torch.int64"), _FakeTensor(good, cuda=False),
                _FakeTensor(good, contiguous=False), _FakeTensor(good, device="cuda:1")):
        with pytest.raises(ig.UsageError):
            ig._check_device_jk(5, _FakeTensor(good), bad, None)
    with pytest.raises(ig.UsageError):
        ig._check_device_jk(5, _FakeTensor(good, dtype="torch.float32"), _FakeTensor(good))


def test_density_count_limits_through_the_c_abi_before_any_device_work():
    """direct.rs:141-148 through the C entry points: the density count is
    checked before the contractor is touched (a dummy handle is never
    dereferenced), with the reference's two panic categories as codes."""
    lib = ig.library()
    fake = ctypes.c_void_p(256)
    for fn in (lib.ig_direct_jk_device, ):
        assert fn(fake, fake, 0, 0.0, 1, fake, fake, None) == ig.E_NODENS
        assert fn(fake, fake, 65, 0.0, 1, fake, fake, None) == ig.E_TOOMANY
    assert lib.ig_direct_jk_device_ws(fake, None, fake, 0, 0.0, 1, fake, fake, 1, None) == ig.E_NODENS
    assert lib.ig_direct_jk_device_ws(fake, None, fake, 65, 0.0, 1, fake, fake, 1, None) == ig.E_TOOMANY
    assert lib.ig_direct_jk_device_noalloc(fake, fake, 0, 0.0, 1, fake, fake, 1, None) == ig.E_NODENS
    assert lib.ig_direct_jk_device_noalloc(fake, fake, 65, 0.0, 1, fake, fake, 1,
                                           None) == ig.E_TOOMANY
    # want_exchange without a K buffer, and a null contractor
    assert lib.ig_direct_jk_device_ws(fake, None, fake, 1, 0.0, 1, fake, None, 1, None) == ig.E_ARG
    assert lib.ig_direct_jk_device_ws(None, None, fake, 1, 0.0, 1, fake, fake, 1, None) == ig.E_ARG
    assert lib.ig_direct_workspace_create(None, ctypes.byref(ctypes.c_void_p())) == ig.E_ARG
    assert lib.ig_grad_workspace_create(None, ctypes.byref(ctypes.c_void_p())) == ig.E_ARG
