"""GPU grouping of raw integer keys (fehip_group; FixedEffect(keys...),
src/FixedEffect.jl:18-21) against numpy's unique on the host: the level
numbering (ascending lexicographic key order) is integer work, so the bar is
bit-exact refs and level counts.  GroupedArray's own numbering for
sort = nothing is unspecified; test/solve.jl:18-24 pins only that residuals do
not depend on it, which the last test checks on the GPU path.
"""
import sys

import numpy as np
import pytest

from fixedeffects_rocm import api

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _host(*keys):
    fe = api.FixedEffect(*keys)
    return fe.refs, fe.n


def _check(*keys):
    refs, n = api.group(*keys)
    want, wn = _host(*keys)
    assert n == wn
    np.testing.assert_array_equal(refs, want)


@pytest.mark.parametrize("N,card,seed", [(1, 1, 1), (1000, 7, 2), (100_000, 5000, 3), (2_000_000, 300_000, 4)])
def test_group_one_column(gpu, N, card, seed):
    _say(f"case: group one column N={N} card={card}")
    rng = np.random.default_rng(seed)
    _check(rng.integers(-card, card, N))


def test_group_extreme_keys(gpu):
    _say("case: group extreme keys")
    rng = np.random.default_rng(5)
    lo, hi = np.iinfo(np.int64).min, np.iinfo(np.int64).max
    k = rng.choice(np.array([lo, hi, 0, -1, 1, lo + 1, hi - 1, 2 ** 40, -(2 ** 40)], np.int64), 50_000)
    _check(k)  # full 64-bit range: eight digit passes
    _check(np.full(1234, 42, np.int64))  # one level


def test_group_several_columns(gpu):
    _say("case: group several columns")
    rng = np.random.default_rng(6)
    N = 300_000
    a = rng.integers(0, 1000, N)
    b = rng.integers(-50, 50, N).astype(np.int32)
    c = rng.integers(0, 3, N).astype(np.uint8)
    _check(a, b)
    _check(a, b, c)
    _check(c, a)


def test_group_empty_and_type_errors(gpu):
    _say("case: group edges")
    refs, n = api.group(np.array([], np.int64))
    assert n == 0 and len(refs) == 0
    with pytest.raises(TypeError):
        api.group(np.array([1.5, 2.5]))
    with pytest.raises(api.DimensionMismatch):
        api.group(np.arange(3), np.arange(4))


def test_residuals_do_not_depend_on_numbering(gpu):
    # test/solve.jl:18-24: grouping the same categories differently (here:
    # scrambled key values) leaves the residuals unchanged
    _say("case: group numbering invariance")
    rng = np.random.default_rng(7)
    N = 50_000
    w = rng.integers(0, 4000, N)
    f = rng.integers(0, 300, N)
    y = rng.standard_normal(N)
    fes_a = [api.FixedEffect(w, device=True), api.FixedEffect(f, device=True)]
    perm_w = rng.permutation(4000) * 7 - 9000
    perm_f = rng.permutation(300) + 10 ** 12
    fes_b = [api.FixedEffect(perm_w[w], device=True), api.FixedEffect(perm_f[f], device=True)]
    ra, _, ca = api.solve_residuals(y.copy(), fes_a, double_precision=True, tol=1e-12, maxiter=100_000)
    rb, _, cb = api.solve_residuals(y.copy(), fes_b, double_precision=True, tol=1e-12, maxiter=100_000)
    assert ca and cb
    assert np.linalg.norm(ra - rb) / np.linalg.norm(ra) <= 1e-8
