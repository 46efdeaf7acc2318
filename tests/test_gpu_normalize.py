"""GPU parity of normalize! and full on the device (src/FixedEffect.jl:57-140)
against the oracle's restatement (itself pinned against scipy's connected
components and the zero-mean invariants in test_oracle.py).

Bars: full is a gather -- bit-exact against the oracle's full of the same
level coefficients.  normalize! sums per component in a different order than
the oracle (the reference's own order is that of a hash set), so it is held
to 1e-12 absolute on coefficients of unit size; component structure (which
levels move together, which are left alone) is exact by construction of the
test: a wrong label moves a coefficient by O(1).
"""
import sys

import numpy as np
import pytest

import oracle
from fixedeffects_rocm import api

pytestmark = pytest.mark.gpu

ATOL = 1e-12


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _blocky(seed, N, n, nblocks, spare=3):
    """nblocks disjoint level ranges per FE (>= nblocks components), plus
    `spare` never-observed levels at the end of every FE."""
    rng = np.random.default_rng(seed)
    blk = rng.integers(0, nblocks, N)
    refs = []
    for k in n:
        per = max(1, k // nblocks)
        refs.append((blk * per + rng.integers(0, per, N) + 1).astype(np.int64))
    return refs, [max(1, k // nblocks) * nblocks + spare for k in n]


def _fes(refs, n, z):
    return [api.FixedEffect.from_refs(r, k, zz) for r, k, zz in zip(refs, n, z)]


CASES = {
    # name: (seed, N, levels, blocks, FEs with an interaction)
    "two_fe_one_component": (11, 5000, (300, 40), 1, ()),
    "three_fe_seven_components": (12, 20_000, (700, 140, 35), 7, ()),
    "interaction_skipped": (13, 20_000, (700, 140, 35), 5, (1,)),
    "one_unit_fe_left_alone": (14, 8000, (200, 50, 20), 3, (0, 2)),
    "single_fe": (15, 3000, (120,), 1, ()),
    "four_fe_many_components": (16, 200_000, (20_000, 4000, 1000, 300), 97, ()),
}


@pytest.mark.parametrize("name", list(CASES))
def test_normalize_matches_oracle(gpu, name):
    seed, N, n, nb, inter = CASES[name]
    _say(f"case: normalize {name}")
    refs, nlev = _blocky(seed, N, n, nb)
    rng = np.random.default_rng(seed + 1000)
    z = [rng.uniform(0.5, 1.5, N) if g in inter else None for g in range(len(n))]
    coefs = [rng.standard_normal(k) for k in nlev]
    want = oracle.normalize([c.copy() for c in coefs], refs, nlev, [zz is None for zz in z])
    with api.FixedEffectSolverROCm(_fes(refs, nlev, z), None, "fp64") as s:
        got = s.normalize([c.copy() for c in coefs])
        again = s.normalize([c.copy() for c in coefs])
    for g, (a, b, c0) in enumerate(zip(got, want, coefs)):
        np.testing.assert_allclose(a, b, rtol=0, atol=ATOL, err_msg=f"FE {g}")
        if inter and g in inter:
            np.testing.assert_array_equal(a, c0)  # interacted FEs untouched (:107)
        np.testing.assert_array_equal(a[-3:], c0[-3:])  # unobserved levels untouched
    for a, b in zip(got, again):
        np.testing.assert_array_equal(a, b)  # fixed-order sums: deterministic


def test_solve_coefficients_full_matches_oracle(gpu):
    _say("case: solve_coefficients_full")
    refs, nlev = _blocky(21, 30_000, (1500, 300, 60), 4)
    rng = np.random.default_rng(22)
    w = rng.lognormal(0, 0.5, len(refs[0]))
    y = rng.standard_normal(len(refs[0]))
    with api.FixedEffectSolverROCm(_fes(refs, nlev, [None] * 3), w, "fp64") as s:
        raw, it_raw, _ = s.solve_coefficients_raw(y.copy(), tol=1e-10, maxiter=10_000)
        obs, lev, it, cv = s.solve_coefficients_full(y.copy(), tol=1e-10, maxiter=10_000)
        obs2, it2, cv2 = s.solve_coefficients(y.copy(), tol=1e-10, maxiter=10_000)
    assert cv and cv2 and it == it_raw == it2
    want_lev = oracle.normalize([c.copy() for c in raw], refs, nlev, [True] * 3)
    for a, b in zip(lev, want_lev):
        np.testing.assert_allclose(a, b, rtol=0, atol=ATOL * max(1.0, np.abs(b).max()))
    want_obs = oracle.full(lev, refs, nlev)
    for a, b, c in zip(obs, want_obs, obs2):
        np.testing.assert_array_equal(a, b)  # full is a gather
        np.testing.assert_array_equal(a, c)
    # fitted values agree with the oracle's own solve (normalization-free)
    o = oracle.OracleSolver(refs, nlev, [None] * 3, w)
    co, io, cvo = o.solve_coefficients(y.copy(), tol=1e-10, maxiter=10_000)
    assert cvo
    fit_o, fit_g = sum(co), sum(obs)
    assert np.linalg.norm(fit_g - fit_o) / np.linalg.norm(fit_o) <= 1e-8


def test_solve_coefficients_full_optional_outputs(gpu):
    _say("case: solve_coefficients_full optional outputs")
    refs, nlev = _blocky(31, 4000, (200, 30), 2)
    y = np.random.default_rng(32).standard_normal(4000)
    with api.FixedEffectSolverROCm(_fes(refs, nlev, [None, None]), None, "fp64") as s:
        obs, lev, _, _ = s.solve_coefficients_full(y.copy(), levels=False)
        obs_b, lev_b, _, _ = s.solve_coefficients_full(y.copy(), per_obs=False)
    assert lev is None and obs_b is None
    for a, b in zip(obs, oracle.full(lev_b, refs, nlev)):
        np.testing.assert_array_equal(a, b)
