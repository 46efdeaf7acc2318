"""GPU checks of the FixedEffectModels-style caller (partial_out / fe_ols):
coefficients equal those of the dense dummy-variable regression, on the
estimation sample, with weights and an interaction."""
import sys

import numpy as np
import pytest

from fixedeffects_rocm import api, models

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _dummies(refs, n, z=None):
    D = np.zeros((len(refs), n))
    D[np.arange(len(refs)), refs - 1] = 1.0 if z is None else z
    return D


def test_fe_ols_matches_dummy_regression(gpu):
    _say("case: fe_ols")
    rng = np.random.default_rng(4)
    N = 3000
    f1 = rng.integers(1, 201, N)
    f2 = rng.integers(1, 31, N)
    z = rng.uniform(0.5, 1.5, N)
    X = rng.standard_normal((N, 3))
    y = X @ np.array([0.5, -1.0, 2.0]) + rng.standard_normal(200)[f1 - 1] + rng.standard_normal(30)[f2 - 1] \
        + 0.3 * z * rng.standard_normal(30)[f2 - 1] + rng.standard_normal(N)
    w = rng.lognormal(0, 0.5, N)
    esample = rng.random(N) < 0.9
    fes = [api.FixedEffect.from_refs(f1, 200), api.FixedEffect.from_refs(f2, 30),
           api.FixedEffect.from_refs(f2, 30, z)]
    beta, resid, iters, conv = models.fe_ols(y, X, fes, w, esample, tol=1e-12)
    assert all(conv)
    # dense regression on the estimation sample with every dummy column
    e = esample
    D = np.column_stack([X[e], _dummies(f1[e], 200), _dummies(f2[e], 30), _dummies(f2[e], 30, z[e])])
    sw = np.sqrt(w[e])
    coef, *_ = np.linalg.lstsq(D * sw[:, None], y[e] * sw, rcond=None)
    _say(f"  beta {beta} dense {coef[:3]} iterations {iters}")
    np.testing.assert_allclose(beta, coef[:3], rtol=1e-8, atol=1e-10)
    # residuals of the full model
    full_resid = y[e] - D @ coef
    np.testing.assert_allclose(resid, full_resid, rtol=1e-7, atol=1e-9)


def test_partial_out_many_columns(gpu):
    # more columns than one batch holds, with the reference's default tol
    _say("case: partial_out 40 columns")
    rng = np.random.default_rng(5)
    N = 20_000
    f1 = rng.integers(1, 2001, N)
    f2 = rng.integers(1, 51, N)
    Y = rng.standard_normal((N, 40)) + rng.standard_normal((2000, 40))[f1 - 1]
    fes = [api.FixedEffect.from_refs(f1, 2000), api.FixedEffect.from_refs(f2, 50)]
    R, iters, conv = models.partial_out(Y, fes, tol=1e-10)
    assert R.shape == Y.shape and all(conv)
    # normal equations: every level's residual sum vanishes
    for f, n in ((f1, 2000), (f2, 50)):
        sums = np.stack([np.bincount(f - 1, weights=R[:, k], minlength=n) for k in range(40)], 1)
        assert np.abs(sums).max() <= 1e-7 * np.abs(Y).sum(0).max()
