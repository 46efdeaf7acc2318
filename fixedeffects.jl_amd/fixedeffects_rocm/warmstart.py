"""Learned warm start for the residualization LSMR.

A small network maps per-level features of a column -- log size, weighted
group mean, degree in the fixed-effect graph, one alternating-projection pass
(all computed on the GPU by fehip_features) -- to a correction of the
alternating-projection coefficients.  LSMR then starts from those
coefficients instead of the reference start x0 = A'b
(src/AbstractFixedEffectSolver.jl:56) and still runs to the same tolerance,
so the result is a solution of the same stopping rule; only the iteration
count changes.

Training data comes from the solver itself: for each training panel, the
target of level k is (b*_k - b_ap_k) / sigma, where b* is the LSMR solution
started from the alternating-projection coefficients b_ap (so targets carry
no arbitrary per-component constants) and sigma the rms of sqrt(w) .* y.
The network is trained on the GPU (fehip_mlp_fit: bf16 MFMA forward and
backward, Adam in FP32).
"""
from __future__ import annotations

import ctypes as C
import os
from typing import List, Optional, Sequence

import numpy as np

from . import api, synth
from ._lib import check, lib

_fp = C.POINTER(C.c_float)
_dp = C.POINTER(C.c_double)
_i64p = C.POINTER(C.c_int64)
_u8p = C.POINTER(C.c_uint8)

N_FEATURES = 16
WEIGHTS_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "warmstart_weights.npz")


class WarmStartNet:
    """The network (fehip_mlp_t): 16 -> 128 -> 128 -> 128 -> 1, ReLU."""

    def __init__(self, seed: int = 0):
        h = C.c_void_p()
        check(lib.fehip_mlp_create(C.byref(h), int(seed)))
        self._h = h
        n = C.c_int64()
        check(lib.fehip_mlp_nparams(self._h, C.byref(n)))
        self.nparams = int(n.value)

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib.fehip_mlp_destroy(self._h)
        self._h = None

    def __del__(self):
        self.close()

    @property
    def params(self) -> np.ndarray:
        out = np.empty(self.nparams, np.float32)
        check(lib.fehip_mlp_get_params(self._h, out.ctypes.data_as(_fp)))
        return out

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.nparams
        check(lib.fehip_mlp_set_params(self._h, p.ctypes.data_as(_fp)))

    def gradient(self, X, t):
        X = np.ascontiguousarray(X, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        g = np.empty(self.nparams, np.float32)
        loss = C.c_double()
        check(lib.fehip_mlp_gradient(self._h, X.ctypes.data_as(_fp), t.ctypes.data_as(_fp), len(t),
                                     g.ctypes.data_as(_fp), C.byref(loss)))
        return float(loss.value), g

    def fit(self, X, t, epochs=5, batch=4096, lr=1e-3, seed=0):
        X = np.ascontiguousarray(X, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        losses = np.zeros(epochs)
        check(lib.fehip_mlp_fit(self._h, X.ctypes.data_as(_fp), t.ctypes.data_as(_fp), len(t), int(epochs),
                                int(batch), float(lr), int(seed), losses.ctypes.data_as(_dp)))
        return losses

    def predict(self, X) -> np.ndarray:
        X = np.ascontiguousarray(X, np.float32)
        y = np.empty(len(X), np.float32)
        check(lib.fehip_mlp_predict(self._h, X.ctypes.data_as(_fp), len(X), y.ctypes.data_as(_fp)))
        return y

    def save(self, path: str = WEIGHTS_PATH, **meta):
        np.savez(path, params=self.params, **{k: np.asarray(v) for k, v in meta.items()})

    @classmethod
    def load(cls, path: str = WEIGHTS_PATH) -> "WarmStartNet":
        d = np.load(path, allow_pickle=False)
        net = cls()
        net.params = d["params"]
        return net


def features(solver: api.FixedEffectSolverROCm, y):
    """(features [p x 16] float32, alternating-projection coefficients [p], sigma)."""
    y = np.ascontiguousarray(y, np.float64)
    F = np.empty((solver.p, N_FEATURES), np.float32)
    ap = np.empty(solver.p, np.float64)
    sig = C.c_double()
    check(lib.fehip_features(solver._h, y.ctypes.data_as(_dp), F.ctypes.data_as(_fp), ap.ctypes.data_as(_dp),
                             C.byref(sig)))
    return F, ap, float(sig.value)


def solve_coefficients_from(solver: api.FixedEffectSolverROCm, y, coef0, tol=1e-10, maxiter=100_000):
    """LSMR from the unscaled coefficients coef0 (p); returns (blocks, iters, converged)."""
    y = np.ascontiguousarray(y, np.float64)
    c0 = np.ascontiguousarray(coef0, np.float64)
    blocks = [np.empty(int(k), np.float64) for k in solver.n]
    arr = (_dp * solver.G)(*[b.ctypes.data_as(_dp) for b in blocks])
    it = C.c_int64()
    cv = C.c_uint8()
    check(lib.fehip_solve_coefficients_x0(solver._h, y.ctypes.data_as(_dp), tol, int(maxiter), c0.ctypes.data_as(_dp),
                                          arr, C.byref(it), C.byref(cv)))
    return blocks, int(it.value), bool(cv.value)


def solve_residuals_from(solver: api.FixedEffectSolverROCm, cols: Sequence[np.ndarray], coef0=None, tol=1e-8,
                         maxiter=10_000):
    """Residualize columns in place from given unscaled start coefficients
    (list of p-vectors, None entries = reference start)."""
    K = len(cols)
    arr = (_dp * K)(*[c.ctypes.data_as(_dp) for c in cols])
    keep = [None if c0 is None else np.ascontiguousarray(c0, np.float64) for c0 in (coef0 or [None] * K)]
    c0arr = (_dp * K)(*[C.cast(None, _dp) if c0 is None else c0.ctypes.data_as(_dp) for c0 in keep])
    it = np.zeros(K, np.int64)
    cv = np.zeros(K, np.uint8)
    check(lib.fehip_solve_residuals_x0(solver._h, arr, K, tol, int(maxiter), c0arr, it.ctypes.data_as(_i64p),
                                       cv.ctypes.data_as(_u8p)))
    return [int(i) for i in it], [bool(c) for c in cv]


def solve_residuals_learned(solver: api.FixedEffectSolverROCm, cols: Sequence[np.ndarray], net: WarmStartNet,
                            tol=1e-8, maxiter=10_000):
    """Residualize columns in place from the network's start."""
    K = len(cols)
    arr = (_dp * K)(*[c.ctypes.data_as(_dp) for c in cols])
    it = np.zeros(K, np.int64)
    cv = np.zeros(K, np.uint8)
    check(lib.fehip_solve_residuals_learned(solver._h, arr, K, tol, int(maxiter), net._h, it.ctypes.data_as(_i64p),
                                            cv.ctypes.data_as(_u8p)))
    return [int(i) for i in it], [bool(c) for c in cv]


def training_rows(panel: synth.Panel, column: int = 0, tol: float = 1e-10):
    """Features and targets of every observed level of one panel column."""
    fes = [api.FixedEffect.from_refs(r, k, z) for r, k, z in zip(panel.refs, panel.n, panel.interactions)]
    y = np.ascontiguousarray(panel.Y[:, column])
    with api.FixedEffectSolverROCm(fes, panel.weights, "fp64") as s:
        F, ap, sig = features(s, y)
        blocks, _, conv = solve_coefficients_from(s, y, ap, tol=tol)
    b = np.concatenate(blocks)
    keep = F[:, 14] == 0.0  # observed levels
    t = ((b - ap) / (sig if sig > 0 else 1.0)).astype(np.float32)
    return F[keep], t[keep], conv


def build_training_set(n_panels: int, seed0: int = 10_000, n_min: float = 1e4, n_max: float = 2e5,
                       log=None):
    Xs, ts = [], []
    for i in range(n_panels):
        p = synth.uneven_panel(seed0 + i, n_min, n_max)
        X, t, conv = training_rows(p, column=int(i % p.K))
        if conv:
            Xs.append(X)
            ts.append(t)
        if log and (i + 1) % 100 == 0:
            log(f"[warmstart] {i + 1}/{n_panels} panels, {sum(len(a) for a in ts)} rows")
    return np.concatenate(Xs), np.concatenate(ts)
