"""CPU-side checks of the boundary: libfehip.so loads and exports every entry
point include/fehip.h declares; argument validation that precedes any device
work; the host mirror's FixedEffect semantics and coefficient post-processing
against the oracle.  No compute call reaches the GPU here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle
from fixedeffects_rocm import _lib
from fixedeffects_rocm import api

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "fehip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fehip_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = C.CDLL(_lib.LIB_PATH)
    declared = _declared()
    assert len(declared) >= 14
    for name in declared:
        assert hasattr(L, name), name
    assert sorted(_lib.EXPORTS) == declared


def test_argument_errors_before_device_work():
    h = C.c_void_p()
    n = np.array([3], np.int64)
    refs = np.array([1, 2, 3], np.int64)
    arr = (C.POINTER(C.c_int64) * 1)(refs.ctypes.data_as(C.POINTER(C.c_int64)))
    rc = _lib.lib.fehip_create(C.byref(h), 0, 1, arr, n.ctypes.data_as(C.POINTER(C.c_int64)), None, None, 0)
    assert rc == _lib.FEHIP_EDIM and b"N must be" in _lib.lib.fehip_last_error()
    rc = _lib.lib.fehip_create(C.byref(h), 3, 0, arr, n.ctypes.data_as(C.POINTER(C.c_int64)), None, None, 0)
    assert rc == _lib.FEHIP_EARG
    rc = _lib.lib.fehip_create(C.byref(h), 3, 1, arr, n.ctypes.data_as(C.POINTER(C.c_int64)), None, None, 7)
    assert rc == _lib.FEHIP_EARG and b"precision" in _lib.lib.fehip_last_error()
    bad = np.array([0], np.int64)
    rc = _lib.lib.fehip_create(C.byref(h), 3, 1, arr, bad.ctypes.data_as(C.POINTER(C.c_int64)), None, None, 0)
    assert rc == _lib.FEHIP_EDIM
    assert _lib.lib.fehip_solve_residuals(None, None, 1, 1e-8, 10, None, None) == _lib.FEHIP_EARG
    _lib.lib.fehip_destroy(None)


def test_fixedeffect_semantics():
    fe = api.FixedEffect(np.array([10, 30, 10, 20]))
    np.testing.assert_array_equal(fe.refs, [1, 3, 1, 2])
    assert fe.n == 3 and fe.interaction is None and len(fe) == 4
    fe2 = api.FixedEffect(np.array([1, 1, 2]), np.array(["a", "b", "a"]))
    assert fe2.n == 3
    with pytest.raises(api.DimensionMismatch):                 # test/types.jl:24
        api.FixedEffect(np.arange(10), interaction=np.ones(5))
    sub = fe[np.array([0, 1])]                                   # test/types.jl:31-44
    np.testing.assert_array_equal(sub.refs, [1, 3])
    assert sub.n == 3 and fe[:] is fe
    with pytest.raises(TypeError):
        fe[np.array([[0, 1]])]


def test_length_mismatch_raises_before_backend():
    fe = api.FixedEffect(np.array([1, 2]))
    with pytest.raises(ValueError, match="FixedEffects must have the same length as y"):   # test/solve.jl:41-42
        api.solve_residuals(np.ones(100), [fe])
    with pytest.raises(ValueError, match="FixedEffects must have the same length as y"):
        api.solve_coefficients(np.ones(100), [fe])


def test_only_rocm_method():
    with pytest.raises(ValueError):
        api.AbstractFixedEffectSolver([api.FixedEffect(np.array([1, 2]))], None, method="cpu")


def test_host_normalize_matches_oracle():
    rng = np.random.default_rng(0)
    N, n = 500, (120, 70, 6)
    refs = [rng.integers(1, k + 1, N).astype(np.int64) for k in n]
    fes = [api.FixedEffect.from_refs(r, k) for r, k in zip(refs, n)]
    coefs = [rng.standard_normal(k) for k in n]
    a = api.normalize([c.copy() for c in coefs], fes)
    b = oracle.normalize([c.copy() for c in coefs], refs, n, [True, True, True])
    for x, y in zip(a, b):
        np.testing.assert_allclose(x, y, rtol=1e-12, atol=1e-12)
    for x, y in zip(api.full(a, fes), oracle.full(b, refs, n)):
        np.testing.assert_allclose(x, y, rtol=1e-12, atol=1e-12)
    # interacted FEs are excluded from normalization
    fes_i = [fes[0], api.FixedEffect.from_refs(refs[1], n[1], rng.uniform(size=N)), fes[2]]
    a = api.normalize([c.copy() for c in coefs], fes_i)
    b = oracle.normalize([c.copy() for c in coefs], refs, n, [True, False, True])
    for x, y in zip(a, b):
        np.testing.assert_allclose(x, y, rtol=1e-12, atol=1e-12)


def test_default_tolerances_per_precision():
    """Inner tol default sqrt(eps(T)) (src/AbstractFixedEffectSolver.jl:49,147)
    with T = Float32 only for the reference Float32 mode; the mixed mode is
    FP64-accurate and keeps Float64's defaults.  Outer defaults (:39-40):
    1e-8 in double precision, 1e-6 otherwise; mixed takes 1e-8."""
    assert api._precision_T("fp64") is np.float64
    assert api._precision_T("mixed") is np.float64
    assert api._precision_T("fp32") is np.float32
    assert api._inner_default_tol(api._precision_T("fp64"), None) == pytest.approx(np.sqrt(2.0 ** -52))
    assert api._inner_default_tol(api._precision_T("mixed"), None) == pytest.approx(np.sqrt(2.0 ** -52))
    assert api._inner_default_tol(api._precision_T("fp32"), None) == pytest.approx(np.sqrt(2.0 ** -23))
    assert api._inner_default_tol(np.float32, 1e-3) == 1e-3
    # outer: method :rocm defaults to Float32 semantics unless told otherwise
    assert api._outer("rocm", None, None) == (False, 1e-6)
    assert api._outer("rocm", True, None) == (True, 1e-8)
    assert api._outer("rocm", None, None, "mixed") == (False, 1e-8)
    assert api._outer("rocm", None, None, "fp64") == (False, 1e-8)
    assert api._outer("rocm", None, None, "fp32") == (False, 1e-6)
    assert api._outer("rocm", None, 1e-4, "mixed") == (False, 1e-4)
    assert api._outer("cpu", None, None) == (True, 1e-8)
