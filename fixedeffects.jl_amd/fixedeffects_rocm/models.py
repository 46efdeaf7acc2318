"""The FixedEffectModels-style caller of the residualization path.

FixedEffectModels.jl partials the fixed effects out of every variable of a
regression (outcome, regressors, instruments) in one call of the multi-column
``solve_residuals!(cols, feM)`` (src/AbstractFixedEffectSolver.jl:74-97) on a
solver built for the estimation sample (``fes[esample]``,
src/FixedEffect.jl:44-50), then runs least squares on the residualized
columns (Frisch-Waugh-Lovell).  ``partial_out`` is that first step on the
GPU (all columns in batched launches of up to 32); ``fe_ols`` adds the
weighted least-squares fit of the second.
"""
from __future__ import annotations

from typing import Optional, Sequence

import numpy as np

from . import api


def partial_out(Y, fes: Sequence[api.FixedEffect], weights=None, esample=None, *, precision: str = "fp64",
                tol: float = 1e-8, maxiter: int = 10_000, algorithm: str = "lsmr"):
    """Residualize every column of Y (N x K) on the fixed effects.

    esample: boolean mask or index array of the observations kept (rows of
    Y, refs, interactions and weights are subset with it, levels keep their
    numbering as in getindex(fe, esample)).  Returns (residuals (n x K,
    float64), iterations per column, converged per column)."""
    Y = np.asarray(Y, np.float64)
    # algorithm="projected": LSMR with the FE of most levels eliminated (the
    # same residuals, fewer iterations)
    if Y.ndim == 1:
        Y = Y[:, None]
    if esample is not None:
        idx = np.asarray(esample)
        Y = Y[idx]
        fes = [fe[idx] for fe in fes]
        if weights is not None:
            weights = np.asarray(weights, np.float64)[idx]
    cols = [np.ascontiguousarray(Y[:, k]) for k in range(Y.shape[1])]
    with api.FixedEffectSolverROCm(list(fes), weights, precision, algorithm=algorithm) as s:
        _, iters, conv = s.solve_residuals_many(cols, tol=tol, maxiter=maxiter)
    return np.stack(cols, axis=1), iters, conv


def fe_ols(y, X, fes: Sequence[api.FixedEffect], weights=None, esample=None, *, precision: str = "fp64",
           tol: float = 1e-8, maxiter: int = 10_000, algorithm: str = "lsmr"):
    """Least squares of y on X with the fixed effects absorbed: partial the
    fixed effects out of [y X] on the GPU, then solve the weighted normal
    equations of the residualized columns.  Returns (coefficients,
    residuals, iterations, converged)."""
    y = np.asarray(y, np.float64)
    X = np.asarray(X, np.float64)
    if X.ndim == 1:
        X = X[:, None]
    R, iters, conv = partial_out(np.column_stack([y, X]), fes, weights, esample, precision=precision, tol=tol,
                                 maxiter=maxiter, algorithm=algorithm)
    w = None
    if weights is not None:
        w = np.asarray(weights, np.float64)
        if esample is not None:
            w = w[np.asarray(esample)]
    sw = np.sqrt(w) if w is not None else 1.0
    ry, rX = R[:, 0], R[:, 1:]
    beta, *_ = np.linalg.lstsq(rX * (sw[:, None] if w is not None else 1.0), ry * sw, rcond=None)
    return beta, ry - rX @ beta, iters, conv
