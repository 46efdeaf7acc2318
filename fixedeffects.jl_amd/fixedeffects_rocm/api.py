"""Host-side mirror of FixedEffects.jl's public API for the ``:rocm`` backend.

Names, argument meanings, defaults and error behaviour follow the reference:

* ``FixedEffect``                    -- src/FixedEffect.jl:7-21, :44-50
* ``AbstractFixedEffectSolver``      -- the backend constructor
  ``AbstractFixedEffectSolver{T}(fes, w, Val{method})`` (src/SolverCPU.jl:55)
* ``FixedEffectSolverROCm.update_weights`` -- update_weights! (src/SolverCPU.jl:67-76)
* ``solve_residuals``                -- src/AbstractFixedEffectSolver.jl:37-68, :74-97
* ``solve_coefficients``             -- src/AbstractFixedEffectSolver.jl:137-160
* ``normalize``, ``components``, ``full`` -- src/FixedEffect.jl:57-140

All arithmetic of the solve runs in libfehip.so on the GPU; there is no CPU
path in this package.
"""
from __future__ import annotations

import ctypes as C
import time
from typing import Iterable, List, Optional, Sequence

import numpy as np

from ._lib import (FEHIP_FP32, FEHIP_FP64, FEHIP_MIXED, check, lib)

_i64p = C.POINTER(C.c_int64)
_i32p = C.POINTER(C.c_int32)
_dp = C.POINTER(C.c_double)
_u8p = C.POINTER(C.c_uint8)


class DimensionMismatch(ValueError):
    """Julia's DimensionMismatch (src/FixedEffect.jl:12-13)."""


def _ptr(a: np.ndarray, ct):
    return a.ctypes.data_as(C.POINTER(ct))


# ---------------------------------------------------------------------------
# FixedEffect
# ---------------------------------------------------------------------------
class FixedEffect:
    """A categorical variable: 1-based ``refs``, optional ``interaction``, ``n`` levels.

    ``FixedEffect(keys1, keys2, ..., interaction=None)`` groups the key arrays
    into dense 1-based refs (the reference uses GroupedArray with
    ``sort = nothing``, src/FixedEffect.jl:18-21; here levels are numbered in
    sorted key order -- level numbering does not change residuals, which is
    what test/solve.jl:18-24 checks).  ``device=True`` numbers integer keys
    on the GPU instead (``group``), with the same numbering.
    ``FixedEffect.from_refs(refs, n, interaction)`` is the inner constructor
    (src/FixedEffect.jl:11-15).
    """

    __slots__ = ("refs", "interaction", "n")

    def __init__(self, *keys, interaction=None, device: bool = False):
        if len(keys) == 0:
            raise ValueError("FixedEffect needs at least one key vector")
        arrs = [np.asarray(k) for k in keys]
        N = len(arrs[0])
        if any(len(a) != N for a in arrs):
            raise DimensionMismatch("all key vectors must have the same length")
        if device:
            refs, n = group(*arrs)
            self._set(refs, interaction, n)
            return
        if len(arrs) == 1:
            _, inv = np.unique(arrs[0], return_inverse=True)
        else:
            rec = np.rec.fromarrays(arrs)
            _, inv = np.unique(rec, return_inverse=True)
        refs = inv.astype(np.int64).reshape(-1) + 1
        n = int(refs.max()) if N else 0
        self._set(refs, interaction, n)

    @classmethod
    def from_refs(cls, refs, n: int, interaction=None) -> "FixedEffect":
        fe = cls.__new__(cls)
        fe._set(np.ascontiguousarray(refs, dtype=np.int64), interaction, int(n))
        return fe

    def _set(self, refs, interaction, n):
        if interaction is not None:
            interaction = np.ascontiguousarray(interaction, dtype=np.float64)
            if len(interaction) != len(refs):
                raise DimensionMismatch(
                    f"cannot match refs of length {len(refs)} with interaction of length {len(interaction)}")
        self.refs = refs
        self.interaction = interaction  # None = UnitWeights
        self.n = n

    def __len__(self) -> int:
        return len(self.refs)

    def __getitem__(self, esample) -> "FixedEffect":
        """getindex(fe, esample) (src/FixedEffect.jl:42-50): keeps n."""
        if isinstance(esample, slice) and esample == slice(None):
            return self
        idx = np.asarray(esample)
        if idx.ndim != 1:
            raise TypeError("esample must be one-dimensional")
        z = None if self.interaction is None else self.interaction[idx]
        return FixedEffect.from_refs(self.refs[idx], self.n, z)

    def __repr__(self) -> str:
        return "Fixed Effects"


def group(*keys):
    """Dense 1-based refs of the distinct combinations of integer key columns,
    computed on the GPU (fehip_group; GroupedArray in src/FixedEffect.jl:18-21),
    numbered in ascending lexicographic key order.  Returns (refs, n)."""
    if len(keys) == 0:
        raise ValueError("group needs at least one key vector")
    cols = []
    for k in keys:
        a = np.asarray(k)
        if a.dtype.kind not in "iub":
            raise TypeError("device grouping takes integer keys")
        if a.dtype.kind == "u" and a.dtype.itemsize == 8 and len(a) and a.max() > np.iinfo(np.int64).max:
            raise TypeError("uint64 keys beyond the int64 range are not supported")
        cols.append(np.ascontiguousarray(a, dtype=np.int64))
    N = len(cols[0])
    if any(len(c) != N for c in cols):
        raise DimensionMismatch("all key vectors must have the same length")
    refs = np.empty(N, np.int64)
    ng = C.c_int64()
    arr = (_i64p * len(cols))(*[_ptr(c, C.c_int64) for c in cols])
    check(lib.fehip_group(N, len(cols), arr, _ptr(refs, C.c_int64), C.byref(ng)))
    return refs, int(ng.value)


# ---------------------------------------------------------------------------
# solver (the :rocm backend object)
# ---------------------------------------------------------------------------
_PRECISIONS = {"fp64": FEHIP_FP64, "fp32": FEHIP_FP32, "mixed": FEHIP_MIXED}
_ALGORITHMS = {"lsmr": 0, "projected": 1}


class FixedEffectSolverROCm:
    """``FixedEffectSolverROCm{T} <: AbstractFixedEffectSolver{T}``.

    Holds a libfehip handle: refs, CSR bucketings, scales and caches live in
    device memory for the life of the object.
    """

    def __init__(self, fes: Sequence[FixedEffect], weights=None, precision: str = "fp64", algorithm: str = "lsmr"):
        fes = list(fes)
        if not fes:
            raise ValueError("at least one FixedEffect is required")
        N = len(fes[0])
        if any(len(fe) != N for fe in fes):
            raise ValueError("FixedEffects must have the same length as y")
        if precision not in _PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(_PRECISIONS)}")
        self.fes = fes
        self.N = N
        self.G = len(fes)
        self.n = np.array([fe.n for fe in fes], dtype=np.int64)
        self.p = int(self.n.sum())
        self.precision = precision
        # element type T of the solver: Float32 only for the reference Float32
        # mode; the mixed mode returns FP64-accurate results (FP64 recurrences
        # and refinement), so it keeps Float64's defaults
        self.T = _precision_T(precision)
        self._refs = [np.ascontiguousarray(fe.refs, dtype=np.int64) for fe in fes]
        self._z = [fe.interaction for fe in fes]
        self.weights = None
        w = self._weights_array(weights)
        refs_arr = (_i64p * self.G)(*[_ptr(r, C.c_int64) for r in self._refs])
        z_arr = (_dp * self.G)(*[C.cast(None, _dp) if z is None else _ptr(z, C.c_double) for z in self._z])
        h = C.c_void_p()
        check(lib.fehip_create(C.byref(h), N, self.G, refs_arr, _ptr(self.n, C.c_int64), z_arr,
                               C.cast(None, _dp) if w is None else _ptr(w, C.c_double), _PRECISIONS[precision]))
        self._h = h
        self.weights = w
        self.algorithm = "lsmr"
        if algorithm != "lsmr":
            self.set_algorithm(algorithm)

    def _weights_array(self, weights):
        if weights is None:
            return None
        w = np.ascontiguousarray(weights, dtype=np.float64)
        if len(w) != self.N:
            raise ValueError("weights must have the same length as the FixedEffects")
        if np.all(w == 1.0):
            # UnitWeights and an all-ones Weights vector give the same arithmetic
            return None
        return w

    # -- lifetime ---------------------------------------------------------
    def close(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib.fehip_destroy(h)
        self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- reference methods -------------------------------------------------
    def update_weights(self, weights) -> "FixedEffectSolverROCm":
        """update_weights!(feM, w) (src/SolverCPU.jl:67-76)."""
        w = self._weights_array(weights)
        check(lib.fehip_update_weights(self._h, C.cast(None, _dp) if w is None else _ptr(w, C.c_double)))
        self.weights = w
        return self

    def _default_tol(self, tol):
        return _inner_default_tol(self.T, tol)

    def solve_residuals(self, y, tol: Optional[float] = None, maxiter: int = 100_000):
        """Inner solve_residuals!(r, feM; tol, maxiter) (:49-68) on one vector,
        overwritten in place; returns (r, iterations, converged)."""
        y = _as_inplace_vector(y, self.N)
        col = (_dp * 1)(_ptr(y, C.c_double))
        it = np.zeros(1, np.int64)
        cv = np.zeros(1, np.uint8)
        check(lib.fehip_solve_residuals(self._h, col, 1, self._default_tol(tol), int(maxiter), _ptr(it, C.c_int64),
                                        _ptr(cv, C.c_uint8)))
        return y, int(it[0]), bool(cv[0])

    def solve_residuals_many(self, xs, tol: Optional[float] = None, maxiter: int = 100_000,
                             progress_bar: bool = False):
        """Multi-column solve_residuals!(xs, feM) (:74-97): all columns in one
        batched solve (up to 32 per launch); returns (xs, iterations, convergeds)."""
        cols = _columns(xs, self.N)
        K = len(cols)
        if K == 0:
            return xs, [], []
        t0 = time.time()
        arr = (_dp * K)(*[_ptr(c, C.c_double) for c in cols])
        it = np.zeros(K, np.int64)
        cv = np.zeros(K, np.uint8)
        check(lib.fehip_solve_residuals(self._h, arr, K, self._default_tol(tol), int(maxiter), _ptr(it, C.c_int64),
                                        _ptr(cv, C.c_uint8)))
        if progress_bar and time.time() - t0 > 2:
            print(f"Demean Variables: {K}/{K}")
        _write_back(xs, cols)
        return xs, [int(i) for i in it], [bool(c) for c in cv]

    def solve_residuals_device(self, dptr: int, K: int, ld: int, tol: Optional[float] = None,
                               maxiter: int = 100_000):
        """Residualize K device-resident columns (column c at dptr + 8*c*ld)."""
        it = np.zeros(K, np.int64)
        cv = np.zeros(K, np.uint8)
        check(lib.fehip_solve_residuals_device(self._h, C.c_void_p(int(dptr)), int(K), int(ld),
                                               self._default_tol(tol), int(maxiter), _ptr(it, C.c_int64),
                                               _ptr(cv, C.c_uint8)))
        return [int(i) for i in it], [bool(c) for c in cv]

    def solve_coefficients(self, y, tol: Optional[float] = None, maxiter: int = 100_000):
        """Inner solve_coefficients!(r, feM; tol, maxiter) (:147-160): returns
        (G per-observation coefficient vectors, div(iterations, 2), converged).
        normalize! and full run on the device (fehip_solve_coefficients_full)."""
        obs, _, it, cv = self.solve_coefficients_full(y, tol, maxiter, levels=False)
        return obs, it, cv

    def solve_coefficients_full(self, y, tol: Optional[float] = None, maxiter: int = 100_000, levels: bool = True,
                                per_obs: bool = True):
        """(per-observation vectors or None, normalized level blocks or None,
        div(iterations, 2), converged)."""
        y = np.ascontiguousarray(y, dtype=np.float64)
        if len(y) != self.N:
            raise ValueError("FixedEffects must have the same length as y")
        blocks = [np.empty(int(k), np.float64) for k in self.n] if levels else None
        obs = [np.empty(self.N, np.float64) for _ in range(self.G)] if per_obs else None
        barr = (_dp * self.G)(*[_ptr(b, C.c_double) for b in blocks]) if levels else None
        oarr = (_dp * self.G)(*[_ptr(o, C.c_double) for o in obs]) if per_obs else None
        it = C.c_int64()
        cv = C.c_uint8()
        check(lib.fehip_solve_coefficients_full(self._h, _ptr(y, C.c_double), self._default_tol(tol), int(maxiter),
                                                barr, oarr, C.byref(it), C.byref(cv)))
        return obs, blocks, int(it.value), bool(cv.value)

    def normalize(self, blocks: List[np.ndarray]) -> List[np.ndarray]:
        """normalize! of G level blocks in place, on the device."""
        if len(blocks) != self.G or any(len(b) != k for b, k in zip(blocks, self.n)):
            raise ValueError("one block of nlevels doubles per fixed effect is required")
        for b in blocks:
            if b.dtype != np.float64 or not b.flags.c_contiguous:
                raise TypeError("blocks must be contiguous float64 arrays")
        arr = (_dp * self.G)(*[_ptr(b, C.c_double) for b in blocks])
        check(lib.fehip_normalize(self._h, arr))
        return blocks

    def solve_coefficients_raw(self, y, tol: Optional[float] = None, maxiter: int = 100_000):
        """LSMR from zero, x .*= scale (:149-158), before normalize!/full."""
        y = np.ascontiguousarray(y, dtype=np.float64)
        if len(y) != self.N:
            raise ValueError("FixedEffects must have the same length as y")
        blocks = [np.empty(int(k), np.float64) for k in self.n]
        arr = (_dp * self.G)(*[_ptr(b, C.c_double) for b in blocks])
        it = C.c_int64()
        cv = C.c_uint8()
        check(lib.fehip_solve_coefficients(self._h, _ptr(y, C.c_double), self._default_tol(tol), int(maxiter), arr,
                                           C.byref(it), C.byref(cv)))
        return blocks, int(it.value), bool(cv.value)

    # -- inspection (test hooks of the C-ABI) ------------------------------
    def csr(self, g: int):
        """CSR bucketing of FE g (0-based perm and offsets)."""
        perm = np.empty(self.N, np.int32)
        off = np.empty(int(self.n[g]) + 1, np.int32)
        check(lib.fehip_get_csr(self._h, g, _ptr(perm, C.c_int32), _ptr(off, C.c_int32)))
        return perm, off

    def scale(self, g: int) -> np.ndarray:
        out = np.empty(int(self.n[g]), np.float64)
        check(lib.fehip_get_scale(self._h, g, _ptr(out, C.c_double)))
        return out

    def cache(self, g: int) -> np.ndarray:
        out = np.empty(self.N, np.float64)
        check(lib.fehip_get_cache(self._h, g, _ptr(out, C.c_double)))
        return out

    def set_algorithm(self, algorithm: str) -> "FixedEffectSolverROCm":
        """"lsmr" (the reference's LSMR on all FEs) or "projected" (LSMR with
        the FE of most levels eliminated: the same least-squares residual and,
        after normalize!, coefficients; fewer iterations)."""
        if algorithm not in _ALGORITHMS:
            raise ValueError(f"algorithm must be one of {sorted(_ALGORITHMS)}")
        check(lib.fehip_set_algorithm(self._h, _ALGORITHMS[algorithm]))
        self.algorithm = algorithm
        return self

    def set_schedule(self, tiles_s: int = 0, tiles_c: int = 0, fuse_dense: int = -1, poll_iters: int = 0):
        """Schedule of the LSMR launches (0 / -1 = defaults); see fehip_set_schedule."""
        check(lib.fehip_set_schedule(self._h, int(tiles_s), int(tiles_c), int(fuse_dense), int(poll_iters)))

    def autotune(self, K: int = 1, reps: int = 10) -> dict:
        """Time candidate schedules on this handle's data and keep the fastest."""
        out = np.zeros(6, np.float64)
        check(lib.fehip_autotune(self._h, int(K), int(reps), _ptr(out, C.c_double)))
        return {"tiles_s": int(out[0]), "tiles_c": int(out[1]), "fuse_dense": int(out[2]),
                "poll_iters": int(out[3]), "ms_per_iter": float(out[4]), "default_ms_per_iter": float(out[5])}

    def plan_info(self, K: int = 1, fp32_storage: bool = False) -> dict:
        """Launch plan of K columns (test hook fehip_plan_info)."""
        out = np.zeros(9, np.int64)
        check(lib.fehip_plan_info(self._h, int(K), 1 if fp32_storage else 0, _ptr(out, C.c_int64)))
        keys = ("fused_dense", "tile_positions", "fused_tiles", "dtx_tiles", "fixup_items", "dense_slots",
                "pieces", "csr_blocks", "fused_cols_per_lane")
        return {k: int(v) for k, v in zip(keys, out)}

    def lsmr_info(self, col: int = 0) -> dict:
        """LSMR scalars of column col of the last solve."""
        out = np.zeros(9, np.float64)
        check(lib.fehip_lsmr_info(self._h, col, _ptr(out, C.c_double)))
        keys = ("iters", "istop", "normr", "normAr", "normA", "condA", "normx", "alpha", "beta")
        d = dict(zip(keys, out.tolist()))
        d["iters"] = int(d["iters"])
        d["istop"] = int(d["istop"])
        return d

    def mul_A(self, x, y=None, a: float = 1.0, b: float = 0.0) -> np.ndarray:
        """mul!(y, A, x, a, b) (src/AbstractFixedEffectLinearMap.jl:37-44)."""
        x = np.ascontiguousarray(x, np.float64)
        y = np.zeros(self.N, np.float64) if y is None else np.array(y, np.float64, copy=True)
        check(lib.fehip_mul_A(self._h, _ptr(y, C.c_double), _ptr(x, C.c_double), a, b))
        return y

    def mul_At(self, y, x=None, a: float = 1.0, b: float = 0.0) -> np.ndarray:
        """mul!(x, A', y, a, b) (src/AbstractFixedEffectLinearMap.jl:26-35)."""
        y = np.ascontiguousarray(y, np.float64)
        x = np.zeros(self.p, np.float64) if x is None else np.array(x, np.float64, copy=True)
        check(lib.fehip_mul_At(self._h, _ptr(x, C.c_double), _ptr(y, C.c_double), a, b))
        return x


def _as_inplace_vector(y, N):
    if not (isinstance(y, np.ndarray) and y.dtype == np.float64 and y.ndim == 1 and y.flags.c_contiguous):
        raise TypeError("y must be a contiguous 1-D float64 numpy array (it is overwritten in place)")
    if len(y) != N:
        raise ValueError("FixedEffects must have the same length as y")
    return y


def _columns(xs, N) -> List[np.ndarray]:
    if isinstance(xs, np.ndarray) and xs.ndim == 2:
        cols = [np.ascontiguousarray(xs[:, j], dtype=np.float64) for j in range(xs.shape[1])]
    else:
        cols = [_as_inplace_vector(x, N) for x in xs]
    for c in cols:
        if len(c) != N:
            raise ValueError("FixedEffects must have the same length as y")
    return cols


def _write_back(xs, cols):
    if isinstance(xs, np.ndarray) and xs.ndim == 2:
        for j, c in enumerate(cols):
            xs[:, j] = c


def AbstractFixedEffectSolver(fes, weights=None, method: str = "rocm", double_precision: bool = True,
                              precision: Optional[str] = None, algorithm: str = "lsmr") -> FixedEffectSolverROCm:
    """Backend constructor AbstractFixedEffectSolver{T}(fes, w, Val{method})
    (src/SolverCPU.jl:55, src/AbstractFixedEffectSolver.jl:43).  Only
    method = "rocm" exists here; precision="mixed" selects the FP32-stream /
    FP64-recurrence mode and algorithm="projected" the LSMR with the FE of
    most levels eliminated (neither is a reference option)."""
    if method != "rocm":
        raise ValueError(f"method {method!r} is not provided by this package (method = 'rocm')")
    if precision is None:
        precision = "fp64" if double_precision else "fp32"
    return FixedEffectSolverROCm(fes, weights, precision, algorithm=algorithm)


def _precision_T(precision: str):
    """Element type of a solver of this precision: Float32 for "fp32" (the
    reference's double_precision = false), Float64 for "fp64" and "mixed"."""
    return np.float32 if precision == "fp32" else np.float64


def _inner_default_tol(T, tol):
    """tol default of the inner solve_residuals!/solve_coefficients!:
    sqrt(eps(T)) (src/AbstractFixedEffectSolver.jl:49, :147)."""
    if tol is None:
        return float(np.sqrt(np.finfo(T).eps))
    return float(tol)


def _outer(method, double_precision, tol, precision=None):
    """Defaults of the outer keyword methods (src/AbstractFixedEffectSolver.jl
    :37-40): double_precision = (method == :cpu), tol = 1e-8 in double
    precision, else 1e-6.  precision = "mixed" (no reference counterpart)
    computes FP64-accurate results and takes the double-precision tol."""
    if double_precision is None:
        double_precision = method == "cpu"  # :39 -- GPU methods default to Float32
    if tol is None:
        tol = 1e-8 if (double_precision or precision in ("fp64", "mixed")) else 1e-6  # :40
    return double_precision, tol


def solve_residuals_panels(solvers, dptrs, Ks, lds, tol: float = 1e-8, maxiter: int = 10_000):
    """Residualize many panels in one call (fehip_solve_residuals_panels):
    panel i has K[i] device-resident columns at dptrs[i] (column c at
    dptrs[i] + 8*c*lds[i]) and its own solver.  All panels' LSMR iterations
    run concurrently; returns (iterations, converged) per panel."""
    P = len(solvers)
    hs = (C.c_void_p * P)(*[s._h for s in solvers])
    ds = (C.c_void_p * P)(*[C.c_void_p(int(d)) for d in dptrs])
    ks = (C.c_int32 * P)(*[int(k) for k in Ks])
    ls = (C.c_int64 * P)(*[int(x) for x in lds])
    its = [np.zeros(int(k), np.int64) for k in Ks]
    cvs = [np.zeros(int(k), np.uint8) for k in Ks]
    ip = (C.POINTER(C.c_int64) * P)(*[_ptr(a, C.c_int64) for a in its])
    cp = (C.POINTER(C.c_uint8) * P)(*[_ptr(a, C.c_uint8) for a in cvs])
    check(lib.fehip_solve_residuals_panels(hs, P, ds, ks, ls, float(tol), int(maxiter), ip, cp))
    return [[int(v) for v in a] for a in its], [[bool(v) for v in a] for a in cvs]


def solve_residuals_panels_host(solvers, columns, tol: float = 1e-8, maxiter: int = 10_000):
    """Residualize many panels from host columns in one call
    (fehip_solve_residuals_panels_host): columns[i] is a list of contiguous
    float64 vectors of panel i, overwritten in place; the library stages them
    in device memory.  Returns (iterations, converged) per panel."""
    P = len(solvers)
    cols = [[_as_inplace_vector(c, s.N) for c in cs] for s, cs in zip(solvers, columns)]
    hs = (C.c_void_p * P)(*[s._h for s in solvers])
    per = [(_dp * len(cs))(*[_ptr(c, C.c_double) for c in cs]) for cs in cols]
    arr = (C.POINTER(_dp) * P)(*[C.cast(a, C.POINTER(_dp)) for a in per])
    ks = (C.c_int32 * P)(*[len(cs) for cs in cols])
    its = [np.zeros(len(cs), np.int64) for cs in cols]
    cvs = [np.zeros(len(cs), np.uint8) for cs in cols]
    ip = (C.POINTER(C.c_int64) * P)(*[_ptr(a, C.c_int64) for a in its])
    cp = (C.POINTER(C.c_uint8) * P)(*[_ptr(a, C.c_uint8) for a in cvs])
    check(lib.fehip_solve_residuals_panels_host(hs, P, arr, ks, float(tol), int(maxiter), ip, cp))
    return [[int(v) for v in a] for a in its], [[bool(v) for v in a] for a in cvs]


def solve_residuals(y, fes, w=None, *, method: str = "rocm", double_precision: Optional[bool] = None,
                    tol: Optional[float] = None, maxiter: int = 10000, precision: Optional[str] = None,
                    progress_bar: bool = True, algorithm: str = "lsmr"):
    """solve_residuals!(y, fes, w; method, double_precision, tol, maxiter)
    (src/AbstractFixedEffectSolver.jl:37-45), and the solver-object forms
    solve_residuals!(y, feM) / solve_residuals!(xs, feM) (:49-97).

    y is overwritten in place.  Returns (r, iterations, converged), or
    (xs, iterations, convergeds) for a collection / 2-D array of columns."""
    if isinstance(fes, FixedEffectSolverROCm):
        feM = fes
        kw = {} if tol is None else {"tol": tol}
        if _is_single(y):
            return feM.solve_residuals(y, maxiter=maxiter, **kw)
        return feM.solve_residuals_many(y, maxiter=maxiter, progress_bar=progress_bar, **kw)
    fes = list(fes)
    N = len(y) if _is_single(y) else (y.shape[0] if isinstance(y, np.ndarray) else len(y[0]))
    if any(len(fe) != N for fe in fes):
        raise ValueError("FixedEffects must have the same length as y")  # :42
    dp, tol = _outer(method, double_precision, tol, precision)
    with AbstractFixedEffectSolver(fes, w, method, dp, precision, algorithm) as feM:
        if _is_single(y):
            return feM.solve_residuals(y, tol=tol, maxiter=maxiter)
        return feM.solve_residuals_many(y, tol=tol, maxiter=maxiter, progress_bar=progress_bar)


def solve_coefficients(y, fes, w=None, *, method: str = "rocm", double_precision: Optional[bool] = None,
                       tol: Optional[float] = None, maxiter: int = 10000, precision: Optional[str] = None,
                       algorithm: str = "lsmr"):
    """solve_coefficients!(y, fes, w; ...) (src/AbstractFixedEffectSolver.jl:137-145)."""
    if isinstance(fes, FixedEffectSolverROCm):
        kw = {} if tol is None else {"tol": tol}
        return fes.solve_coefficients(y, maxiter=maxiter, **kw)
    fes = list(fes)
    if any(len(fe) != len(y) for fe in fes):
        raise ValueError("FixedEffects must have the same length as y")  # :142
    dp, tol = _outer(method, double_precision, tol, precision)
    with AbstractFixedEffectSolver(fes, w, method, dp, precision, algorithm) as feM:
        return feM.solve_coefficients(y, tol=tol, maxiter=maxiter)


def _is_single(y) -> bool:
    return isinstance(y, np.ndarray) and y.ndim == 1


# ---------------------------------------------------------------------------
# coefficient post-processing (host side, as in the reference)
# ---------------------------------------------------------------------------
def components(fes: Sequence[FixedEffect]):
    """Connected components of the observation/level graph
    (src/FixedEffect.jl:68-98) as per-FE level labels (-1 = level not observed)."""
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import connected_components

    n = [fe.n for fe in fes]
    offs = np.concatenate([[0], np.cumsum(n)])
    tot = int(offs[-1])
    if len(fes) == 1:
        rows = np.asarray(fes[0].refs) - 1
        cols = rows
    else:
        r0 = np.asarray(fes[0].refs, np.int64) - 1
        rows = np.concatenate([r0 for _ in fes[1:]])
        cols = np.concatenate([np.asarray(fe.refs, np.int64) - 1 + offs[j] for j, fe in enumerate(fes) if j > 0])
    A = coo_matrix((np.ones(len(rows), np.int8), (rows, cols)), shape=(tot, tot))
    _, lab = connected_components(A, directed=False)
    seen = np.zeros(tot, bool)
    for j, fe in enumerate(fes):
        seen[np.asarray(fe.refs, np.int64) - 1 + offs[j]] = True
    lab = np.where(seen, lab, -1)
    return [lab[offs[j]:offs[j + 1]] for j in range(len(fes))]


def normalize(fecoefs: List[np.ndarray], fes: Sequence[FixedEffect]) -> List[np.ndarray]:
    """normalize! (src/FixedEffect.jl:106-111) with rescale! (:113-135): within
    each connected component, FEs 2..end (UnitWeights interaction only) are
    demeaned and their means added to the first."""
    idx = [j for j, fe in enumerate(fes) if fe.interaction is None]
    if len(idx) < 2:
        return fecoefs
    sel = [fes[j] for j in idx]
    labels = components(sel)
    ncomp = max(int(lb.max()) for lb in labels) + 1
    m = np.zeros(ncomp)
    for jj in range(len(idx) - 1, 0, -1):
        c = fecoefs[idx[jj]]
        lb = labels[jj]
        ok = lb >= 0
        sums = np.bincount(lb[ok], weights=c[ok], minlength=ncomp)
        cnt = np.bincount(lb[ok], minlength=ncomp)
        mj = np.divide(sums, cnt, out=np.zeros(ncomp), where=cnt > 0)
        c[ok] -= mj[lb[ok]]
        m += mj
    c1 = fecoefs[idx[0]]
    lb = labels[0]
    ok = lb >= 0
    c1[ok] += m[lb[ok]]
    return fecoefs


def full(fecoefs: List[np.ndarray], fes: Sequence[FixedEffect]) -> List[np.ndarray]:
    """full (src/FixedEffect.jl:137-140): per-observation coefficient vectors."""
    return [c[np.asarray(fe.refs, np.int64) - 1] for c, fe in zip(fecoefs, fes)]
