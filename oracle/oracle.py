"""ctypes bindings for liboracle.so (the CPU restatement, test infrastructure).

The outer functions mirror the reference's keyword API
(src/AbstractFixedEffectSolver.jl:37-45 and :137-145).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

_i64p = C.POINTER(C.c_int64)
_i32p = C.POINTER(C.c_int32)
_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_vp = C.c_void_p


class LsmrInfo(C.Structure):
    _fields_ = [
        ("iters", C.c_int64),
        ("istop", C.c_int32),
        ("converged", C.c_int32),
        ("normr", C.c_double),
        ("normAr", C.c_double),
        ("normA", C.c_double),
        ("condA", C.c_double),
        ("normx", C.c_double),
    ]


def build() -> str:
    """Compile liboracle.so with oracle/Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)
    return _LIB_PATH


def _load():
    if not os.path.exists(_LIB_PATH):
        build()
    L = C.CDLL(_LIB_PATH)
    L.oracle_bucketize.argtypes = [_i64p, C.c_int64, C.c_int64, _i64p, _i64p]
    L.oracle_components.argtypes = [C.c_int64, C.c_int32, C.POINTER(_i64p), _i64p, C.POINTER(_i64p)]
    L.oracle_components.restype = C.c_int64
    L.oracle_normalize.argtypes = [_dp, C.c_int64, C.c_int32, C.POINTER(_i64p), _i64p, _i32p]
    L.oracle_full.argtypes = [_dp, C.c_int64, C.c_int32, C.POINTER(_i64p), _i64p, _dp]
    for s, tp in (("d", _dp), ("f", _fp)):
        ct = C.c_double if s == "d" else C.c_float
        getattr(L, f"oracle_scale_{s}").argtypes = [tp, C.c_int64, _i64p, _dp, _dp, C.c_int64]
        getattr(L, f"oracle_cache_{s}").argtypes = [tp, _i64p, _dp, _dp, tp, C.c_int64, C.c_int]
        getattr(L, f"oracle_gather_{s}").argtypes = [tp, _i64p, ct, tp, tp, C.c_int64]
        getattr(L, f"oracle_scatter_{s}").argtypes = [tp, ct, tp, _i64p, tp, C.c_int64, C.c_int]
        f = getattr(L, f"oracle_create_{s}")
        f.argtypes = [C.c_int64, C.c_int32, C.POINTER(_i64p), _i64p, C.POINTER(_dp), _dp, C.c_int]
        f.restype = _vp
        getattr(L, f"oracle_destroy_{s}").argtypes = [_vp]
        getattr(L, f"oracle_update_weights_{s}").argtypes = [_vp, _dp]
        getattr(L, f"oracle_get_scale_{s}").argtypes = [_vp, C.c_int32, tp]
        getattr(L, f"oracle_get_cache_{s}").argtypes = [_vp, C.c_int32, tp]
        getattr(L, f"oracle_mul_A_{s}").argtypes = [_vp, tp, tp, ct, ct]
        getattr(L, f"oracle_mul_At_{s}").argtypes = [_vp, tp, tp, ct, ct]
        getattr(L, f"oracle_lsmr_{s}").argtypes = [_vp, tp, tp, C.c_double, C.c_double, C.c_double,
                                                   C.c_int64, C.POINTER(LsmrInfo)]
        getattr(L, f"oracle_solve_residuals_{s}").argtypes = [_vp, _dp, C.c_double, C.c_int64, _i64p, _i32p]
        getattr(L, f"oracle_solve_coefficients_{s}").argtypes = [_vp, _dp, C.c_double, C.c_int64, _dp,
                                                                 _i64p, _i32p]
    return L


lib = _load()


def _p(a: np.ndarray, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def _refs_arr(refs_list):
    refs = [np.ascontiguousarray(r, dtype=np.int64) for r in refs_list]
    arr = (_i64p * len(refs))(*[_p(r, C.c_int64) for r in refs])
    return refs, arr


def bucketize(refs, n: int):
    """bucketize_refs (ext/MetalExt.jl:40-64): 1-based perm and offsets."""
    refs = np.ascontiguousarray(refs, dtype=np.int64)
    perm = np.empty(len(refs), np.int64)
    offsets = np.empty(n + 1, np.int64)
    lib.oracle_bucketize(_p(refs, C.c_int64), len(refs), n, _p(perm, C.c_int64), _p(offsets, C.c_int64))
    return perm, offsets


def components(refs_list, n_list):
    """components (src/FixedEffect.jl:68-98) as per-FE level labels."""
    refs, arr = _refs_arr(refs_list)
    n = np.asarray(n_list, np.int64)
    labels = [np.empty(int(k), np.int64) for k in n]
    larr = (_i64p * len(labels))(*[_p(lb, C.c_int64) for lb in labels])
    nc = lib.oracle_components(len(refs[0]), len(refs), arr, _p(n, C.c_int64), larr)
    return int(nc), labels


def normalize(coefs, refs_list, n_list, unit_interaction):
    """normalize! (src/FixedEffect.jl:106-111); coefs is a list of G arrays."""
    refs, arr = _refs_arr(refs_list)
    n = np.asarray(n_list, np.int64)
    flat = np.ascontiguousarray(np.concatenate([np.asarray(c, np.float64) for c in coefs]))
    ui = np.asarray([1 if u else 0 for u in unit_interaction], np.int32)
    lib.oracle_normalize(_p(flat, C.c_double), len(refs[0]), len(refs), arr, _p(n, C.c_int64),
                         _p(ui, C.c_int32))
    out, o = [], 0
    for k in n:
        out.append(flat[o:o + k].copy())
        o += k
    return out


def full(coefs, refs_list, n_list):
    """full (src/FixedEffect.jl:137-140)."""
    refs, arr = _refs_arr(refs_list)
    n = np.asarray(n_list, np.int64)
    flat = np.ascontiguousarray(np.concatenate([np.asarray(c, np.float64) for c in coefs]))
    N = len(refs[0])
    out = np.empty(len(refs) * N, np.float64)
    lib.oracle_full(_p(flat, C.c_double), N, len(refs), arr, _p(n, C.c_int64), _p(out, C.c_double))
    return [out[g * N:(g + 1) * N] for g in range(len(refs))]


class OracleSolver:
    """AbstractFixedEffectSolver{T}(fes, w, Val{:cpu}) (src/SolverCPU.jl:55-64)."""

    def __init__(self, refs_list, n_list, interactions=None, weights=None,
                 double_precision: bool = True, nthreads: int = 1):
        self.s = "d" if double_precision else "f"
        self.dtype = np.float64 if double_precision else np.float32
        self.ct = C.c_double if double_precision else C.c_float
        self._refs, arr = _refs_arr(refs_list)
        self.N = len(self._refs[0])
        self.G = len(self._refs)
        self.n = np.asarray(n_list, np.int64)
        self.p = int(self.n.sum())
        for r in self._refs:
            if len(r) != self.N:
                raise ValueError("FixedEffects must have the same length as y")
        zs = interactions if interactions is not None else [None] * self.G
        self._z = [None if z is None else np.ascontiguousarray(z, np.float64) for z in zs]
        zarr = (_dp * self.G)(*[C.cast(None, _dp) if z is None else _p(z, C.c_double) for z in self._z])
        self._w = None if weights is None else np.ascontiguousarray(weights, np.float64)
        wp = C.cast(None, _dp) if self._w is None else _p(self._w, C.c_double)
        self.h = getattr(lib, f"oracle_create_{self.s}")(self.N, self.G, arr, _p(self.n, C.c_int64), zarr, wp,
                                                         int(nthreads))

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            getattr(lib, f"oracle_destroy_{self.s}")(h)
            self.h = None

    def update_weights(self, weights):
        """update_weights! (src/SolverCPU.jl:67-76)."""
        self._w = None if weights is None else np.ascontiguousarray(weights, np.float64)
        wp = C.cast(None, _dp) if self._w is None else _p(self._w, C.c_double)
        getattr(lib, f"oracle_update_weights_{self.s}")(self.h, wp)

    def scale(self, g: int) -> np.ndarray:
        out = np.empty(int(self.n[g]), self.dtype)
        getattr(lib, f"oracle_get_scale_{self.s}")(self.h, g, _p(out, self.ct))
        return out

    def cache(self, g: int) -> np.ndarray:
        out = np.empty(self.N, self.dtype)
        getattr(lib, f"oracle_get_cache_{self.s}")(self.h, g, _p(out, self.ct))
        return out

    def mul_A(self, x, y=None, a=1.0, b=0.0):
        """mul!(y, A, x, a, b) (src/AbstractFixedEffectLinearMap.jl:37-44)."""
        x = np.ascontiguousarray(x, self.dtype)
        y = np.zeros(self.N, self.dtype) if y is None else np.array(y, self.dtype, copy=True)
        getattr(lib, f"oracle_mul_A_{self.s}")(self.h, _p(y, self.ct), _p(x, self.ct), a, b)
        return y

    def mul_At(self, y, x=None, a=1.0, b=0.0):
        """mul!(x, A', y, a, b) (src/AbstractFixedEffectLinearMap.jl:26-35)."""
        y = np.ascontiguousarray(y, self.dtype)
        x = np.zeros(self.p, self.dtype) if x is None else np.array(x, self.dtype, copy=True)
        getattr(lib, f"oracle_mul_At_{self.s}")(self.h, _p(x, self.ct), _p(y, self.ct), a, b)
        return x

    def lsmr(self, x0, b, atol=1e-6, btol=1e-6, conlim=1e8, maxiter=None):
        """lsmr! (src/utils/lsmr.jl:46-222); returns (x, info)."""
        x = np.array(x0, self.dtype, copy=True)
        bb = np.array(b, self.dtype, copy=True)
        if maxiter is None:
            maxiter = max(self.N, self.p)
        info = LsmrInfo()
        getattr(lib, f"oracle_lsmr_{self.s}")(self.h, _p(x, self.ct), _p(bb, self.ct), atol, btol, conlim,
                                              int(maxiter), C.byref(info))
        return x, info

    def solve_residuals(self, y, tol=None, maxiter=100_000):
        """inner solve_residuals! (src/AbstractFixedEffectSolver.jl:49-68)."""
        if tol is None:
            tol = float(np.sqrt(np.finfo(self.dtype).eps))
        r = np.array(y, np.float64, copy=True)
        it = C.c_int64()
        cv = C.c_int32()
        getattr(lib, f"oracle_solve_residuals_{self.s}")(self.h, _p(r, C.c_double), tol, int(maxiter),
                                                         C.byref(it), C.byref(cv))
        return r, int(it.value), bool(cv.value)

    def solve_coefficients_raw(self, y, tol=None, maxiter=100_000):
        """solve_coefficients! up to x .*= scale (:147-158): unnormalized blocks."""
        if tol is None:
            tol = float(np.sqrt(np.finfo(self.dtype).eps))
        y = np.ascontiguousarray(y, np.float64)
        coef = np.empty(self.p, np.float64)
        it = C.c_int64()
        cv = C.c_int32()
        getattr(lib, f"oracle_solve_coefficients_{self.s}")(self.h, _p(y, C.c_double), tol, int(maxiter),
                                                            _p(coef, C.c_double), C.byref(it), C.byref(cv))
        blocks, o = [], 0
        for k in self.n:
            blocks.append(coef[o:o + k].copy())
            o += k
        return blocks, int(it.value), bool(cv.value)

    def solve_coefficients(self, y, tol=None, maxiter=100_000):
        """inner solve_coefficients! (:147-160): normalized, expanded per obs."""
        blocks, it, cv = self.solve_coefficients_raw(y, tol, maxiter)
        unit = [z is None for z in self._z]
        blocks = normalize(blocks, self._refs, self.n, unit)
        return full(blocks, self._refs, self.n), it, cv


def _outer_defaults(method, double_precision, tol):
    if double_precision is None:
        double_precision = method == "cpu"
    if tol is None:
        tol = 1e-8 if double_precision else 1e-6
    return double_precision, tol


def solve_residuals(y, refs_list, n_list, weights=None, interactions=None, *, method="cpu",
                    double_precision=None, tol=None, maxiter=10000, nthreads=1):
    """outer solve_residuals! (src/AbstractFixedEffectSolver.jl:37-45)."""
    if any(len(r) != len(y) for r in refs_list):
        raise ValueError("FixedEffects must have the same length as y")
    dp, tol = _outer_defaults(method, double_precision, tol)
    s = OracleSolver(refs_list, n_list, interactions, weights, dp, nthreads)
    return s.solve_residuals(y, tol=tol, maxiter=maxiter)


def solve_coefficients(y, refs_list, n_list, weights=None, interactions=None, *, method="cpu",
                       double_precision=None, tol=None, maxiter=10000, nthreads=1):
    """outer solve_coefficients! (src/AbstractFixedEffectSolver.jl:137-145)."""
    if any(len(r) != len(y) for r in refs_list):
        raise ValueError("FixedEffects must have the same length as y")
    dp, tol = _outer_defaults(method, double_precision, tol)
    s = OracleSolver(refs_list, n_list, interactions, weights, dp, nthreads)
    return s.solve_coefficients(y, tol=tol, maxiter=maxiter)
