"""GPU parity: libfehip (through its C-ABI, via the host mirror) against the
oracle on the same seeded inputs.

Bars:
* CSR permutation and offsets: bit-exact (integer).
* scale / cache (FP64): bit-exact -- same summation order, same IEEE ops.
* one matvec (FP64): rtol 1e-13 (parallel summation order differs).
* residuals (FP64): at equal LSMR iteration counts (tol = 0, k <= 12) the
  iterates agree to 1e-12; converged (tol = 1e-12) ||r_gpu - r_oracle|| /
  ||r_oracle|| <= 1e-8; at the reference's tol = 1e-8 the iteration counts
  agree within +-2, the GPU residual is at least as close to the exact
  least-squares residual as the oracle's own (within a factor 3), and in FP64
  it lies within a quarter of the oracle's own tolerance-level error of the
  oracle's residual (measured: at most a tenth).  LSMR in
  finite precision loses orthogonality, and a reordering of the oracle's own
  sums already moves its mid-run scalars by up to 1e-2 and its stopping
  iteration by one, so tol-level residual differences are not parity defects.
* reference Float32 mode: Float32.(r) matches at sqrt(eps(Float32)), as
  test/solve.jl:37 checks.
"""
import json
import os
import sys

import numpy as np
import pytest

import oracle
from fixedeffects_rocm import api, synth
from fixedeffects_rocm._lib import FEHIP_EREF, FehipError

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "solve_jl.json")))
RES_RTOL = 1e-8


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _fes(refs, n, z=None):
    z = z or [None] * len(refs)
    return [api.FixedEffect.from_refs(r, k, zz) for r, k, zz in zip(refs, n, z)]


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _rand_problem(seed, N, n, interaction=(), weights=True, empty_level=False):
    rng = np.random.default_rng(seed)
    refs = []
    for g, k in enumerate(n):
        r = rng.integers(1, k + 1, N).astype(np.int64)
        if N >= k:
            r[:k] = np.arange(1, k + 1)
            rng.shuffle(r)
        if empty_level and k > 2:
            r[r == 2] = 1  # level 2 never observed
        refs.append(r)
    z = [rng.uniform(0.5, 1.5, N) if g in interaction else None for g in range(len(n))]
    w = rng.lognormal(0, 0.5, N) if weights else None
    y = rng.standard_normal(N)
    return refs, list(n), z, w, y


# -- golden vectors (test/solve.jl) ---------------------------------------------

def test_golden_fp64(gpu):
    _say("case: golden fp64")
    x = np.array(GOLD["x"])
    fes = [api.FixedEffect(np.array(GOLD["p1"])), api.FixedEffect(np.array(GOLD["p2"]))]
    r, it, cv = api.solve_residuals(x.copy(), fes, double_precision=True)
    assert cv
    np.testing.assert_allclose(r, GOLD["r_ols"], rtol=GOLD["rtol_float64"], atol=0)
    coefs, it, cv = api.solve_coefficients(x.copy(), fes, double_precision=True)
    assert cv
    np.testing.assert_allclose(x - coefs[0] - coefs[1], GOLD["r_ols"], rtol=GOLD["rtol_float64"], atol=0)


def test_golden_fp32_default(gpu):
    # method = :rocm defaults to double_precision = false, tol = 1e-6 (:39-40)
    _say("case: golden fp32")
    x = np.array(GOLD["x"])
    fes = [api.FixedEffect(np.array(GOLD["p1"])), api.FixedEffect(np.array(GOLD["p2"]))]
    r, it, cv = api.solve_residuals(x.copy(), fes)
    a, b = r.astype(np.float32), np.array(GOLD["r_ols"], np.float32)
    assert np.linalg.norm(a - b) <= np.sqrt(np.finfo(np.float32).eps) * np.linalg.norm(b)


# -- CSR bucketing -----------------------------------------------------------------

@pytest.mark.parametrize("N,n,skew", [(1, 1, False), (5, 3, False), (4097, 7, False), (100_000, 100_000, False),
                                      (300_000, 1000, True), (262_144, 65_536, False), (1_000_003, 300_000, True)])
def test_csr_bit_exact(gpu, N, n, skew):
    _say(f"case: csr N={N} n={n} skew={skew}")
    rng = np.random.default_rng(N)
    if skew:
        sizes = synth.zipf_sizes(N, n, 1.1, 1)
        r0 = np.repeat(np.arange(1, n + 1), sizes)
        rng.shuffle(r0)
    else:
        r0 = rng.integers(1, n + 1, N)
    r1 = rng.integers(1, max(2, n // 3) + 1, N)
    refs = [r0.astype(np.int64), r1.astype(np.int64)]
    ns = [n, max(2, n // 3)]
    with api.FixedEffectSolverROCm(_fes(refs, ns), None, "fp64") as s:
        for g in range(2):
            perm, off = s.csr(g)
            operm, ooff = oracle.bucketize(refs[g], ns[g])
            np.testing.assert_array_equal(perm, operm - 1)
            np.testing.assert_array_equal(off, ooff - 1)


def test_bad_refs_rejected(gpu):
    _say("case: bad refs")
    with pytest.raises(FehipError) as ei:
        api.FixedEffectSolverROCm([api.FixedEffect.from_refs(np.array([1, 2, 4]), 3)], None, "fp64")
    assert ei.value.code == FEHIP_EREF
    with pytest.raises(FehipError):
        api.FixedEffectSolverROCm([api.FixedEffect.from_refs(np.array([1, 0, 2]), 3)], None, "fp64")


# -- scale! / cache! -------------------------------------------------------------

@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_scale_cache_exact(gpu, prec):
    _say(f"case: scale/cache {prec}")
    refs, n, z, w, _ = _rand_problem(21, 50_000, (9000, 700, 12), interaction=(2,), empty_level=True)
    o = oracle.OracleSolver(refs, n, z, w, double_precision=(prec == "fp64"))
    with api.FixedEffectSolverROCm(_fes(refs, n, z), w, prec) as s:
        for g in range(3):
            np.testing.assert_array_equal(s.scale(g), o.scale(g).astype(np.float64))
            np.testing.assert_array_equal(s.cache(g), o.cache(g).astype(np.float64))
        # update_weights! recomputes both
        w2 = np.random.default_rng(5).lognormal(0, 1, len(w))
        s.update_weights(w2)
        o.update_weights(w2)
        for g in range(3):
            np.testing.assert_array_equal(s.scale(g), o.scale(g).astype(np.float64))
            np.testing.assert_array_equal(s.cache(g), o.cache(g).astype(np.float64))


# -- one matvec in each direction ---------------------------------------------------

@pytest.mark.parametrize("n", [(5000, 300, 20), (200_000, 40), (64, 3000, 7, 2)])
def test_matvec_fp64(gpu, n):
    _say(f"case: matvec n={n}")
    refs, n, z, w, y = _rand_problem(31, 200_003, n, interaction=(1,), empty_level=True)
    o = oracle.OracleSolver(refs, n, z, w)
    rng = np.random.default_rng(1)
    x = rng.standard_normal(o.p)
    y0 = rng.standard_normal(o.N)
    x0 = rng.standard_normal(o.p)
    with api.FixedEffectSolverROCm(_fes(refs, n, z), w, "fp64") as s:
        for a, b in ((1.0, 0.0), (-1.0, 1.0), (0.37, -2.5)):
            got = s.mul_A(x, y0, a, b)
            np.testing.assert_allclose(got, o.mul_A(x, y0, a, b), rtol=1e-13, atol=1e-13)
            got = s.mul_At(y, x0, a, b)
            np.testing.assert_allclose(got, o.mul_At(y, x0, a, b), rtol=1e-12, atol=1e-12)


# -- full solves ---------------------------------------------------------------------

def _check_solve(refs, n, z, w, Y, prec="fp64", tol=1e-8, iter_slack=2):
    """GPU vs oracle on K columns at `tol`; also against a tight-tolerance
    oracle solve (tol 1e-13) that stands in for the exact residual."""
    K = Y.shape[1]
    o = oracle.OracleSolver(refs, n, z, w, double_precision=(prec == "fp64"), nthreads=8)
    want, tight = [], []
    for c in range(K):
        want.append(o.solve_residuals(Y[:, c].copy(), tol=tol, maxiter=10000))
        tight.append(o.solve_residuals(Y[:, c].copy(), tol=1e-13, maxiter=100000)[0])
    with api.FixedEffectSolverROCm(_fes(refs, n, z), w, prec) as s:
        cols = [Y[:, c].copy() for c in range(K)]
        if K == 1:
            r, it, cv = s.solve_residuals(cols[0], tol=tol, maxiter=10000)
            got = [(r, it, cv)]
        else:
            _, its, cvs = s.solve_residuals_many(cols, tol=tol, maxiter=10000)
            got = list(zip(cols, its, cvs))
    worst = 0.0
    for c, ((rg, ig, cg), (ro, io, co), rt) in enumerate(zip(got, want, tight)):
        worst = max(worst, _rel(rg, ro))
        if c < 4:
            _say(f"  col {c} tol {tol:g}: iters gpu {ig} oracle {io}; rel err vs oracle {_rel(rg, ro):.2e}, "
                 f"to the tight solve gpu {_rel(rg, rt):.2e} oracle {_rel(ro, rt):.2e}")
        assert cg == co
        assert abs(ig - io) <= iter_slack, (ig, io)
        assert _rel(rg, rt) <= 3 * _rel(ro, rt) + 1e-12, (_rel(rg, rt), _rel(ro, rt))
        if prec == "fp64" and tol >= 1e-9:
            # the same answer up to a fraction of the tolerance-level error:
            # measured at most ~0.1 of the oracle's distance to the converged
            # residual, also where the stop is one iteration apart
            assert _rel(rg, ro) <= 0.25 * _rel(ro, rt) + 1e-12, (c, _rel(rg, ro), _rel(ro, rt))
    return worst, [g[1] for g in got], [w_[1] for w_ in want]


def test_solve_cpu_two_way_config(gpu):
    # N=100,000, FE1 uniform 1:10000, FE2 uniform 1:100, unit weights
    _say("case: cpu_two_way full size")
    p = synth.config("cpu_two_way")
    worst, ig, io = _check_solve(p.refs, p.n, p.interactions, p.weights, p.Y)
    _say("  rel err", worst, "iters gpu", ig, "oracle", io)


def test_solve_worker_firm_scaled(gpu):
    # the 1e7 worker/firm/year workload at 1/50 scale: Zipf workers, mobility, weights
    _say("case: worker_firm scale 0.02")
    p = synth.config("worker_firm_1e7", scale=0.02)
    worst, ig, io = _check_solve(p.refs, p.n, p.interactions, p.weights, p.Y)
    _say("  rel err", worst, "iters gpu", ig, "oracle", io)


@pytest.mark.parametrize("name,scale", [("worker_firm_1e7", 0.02), ("cpu_two_way", 1.0)])
def test_solve_converged_parity(gpu, name, scale):
    # both sides converged (tol 1e-12): residuals agree to 1e-8 relative
    _say(f"case: converged parity {name}")
    p = synth.config(name, scale=scale)
    worst, ig, io = _check_solve(p.refs, p.n, p.interactions, p.weights, p.Y, tol=1e-12, iter_slack=10)
    _say("  rel err", worst, "iters gpu", ig, "oracle", io)
    assert worst <= RES_RTOL


def test_solve_interaction_and_single_fe(gpu):
    _say("case: interaction / single FE")
    refs, n, z, w, y = _rand_problem(41, 30_000, (3000, 200, 9), interaction=(2,))
    worst, _, _ = _check_solve(refs, n, z, w, y[:, None])
    worst, _, _ = _check_solve(refs, n, z, w, y[:, None], tol=1e-12, iter_slack=10)
    assert worst <= RES_RTOL
    refs1, n1, z1, w1, y1 = _rand_problem(42, 30_000, (3000,), interaction=(0,))
    worst, ig, _ = _check_solve(refs1, n1, z1, w1, y1[:, None])
    assert worst <= 1e-12 and ig == [1]


def test_solve_batched_columns(gpu):
    _say("case: batched K=5 and K=32")
    p = synth.config("worker_firm_1e7_k32", scale=0.01)
    worst, ig, io = _check_solve(p.refs, p.n, p.interactions, p.weights, p.Y[:, :5])
    _say("  K=5 rel err", worst, ig, io)
    worst, ig, io = _check_solve(p.refs, p.n, p.interactions, p.weights, p.Y, tol=1e-12, iter_slack=10)
    _say("  K=32 converged rel err", worst)
    assert worst <= RES_RTOL


def test_solve_fp32_reference_mode(gpu):
    _say("case: fp32 mode")
    refs, n, z, w, y = _rand_problem(51, 40_000, (4000, 300, 10))
    o = oracle.OracleSolver(refs, n, z, w, double_precision=False)
    ro, io, co = o.solve_residuals(y.copy(), tol=1e-6, maxiter=10000)
    with api.FixedEffectSolverROCm(_fes(refs, n, z), w, "fp32") as s:
        r, it, cv = s.solve_residuals(y.copy(), tol=1e-6, maxiter=10000)
    assert cv == co
    assert _rel(r, ro) <= 1e-4
    assert abs(it - io) <= max(3, io // 10)


def test_mixed_precision_residuals(gpu):
    # FP32 streams + FP64 scalars, then FP64 LSMR iterations: residuals as
    # accurate as an FP64 solve at the same tol, and <= 1e-8 when converged
    _say("case: mixed precision")
    p = synth.config("worker_firm_1e7_k32", scale=0.02)
    Y = p.Y[:, :4]
    o = oracle.OracleSolver(p.refs, p.n, p.interactions, p.weights, nthreads=8)
    with api.FixedEffectSolverROCm(_fes(p.refs, p.n), p.weights, "mixed") as s:
        for tol in (1e-8, 1e-12):
            cols = [Y[:, c].copy() for c in range(4)]
            _, its, cvs = s.solve_residuals_many(cols, tol=tol, maxiter=10000)
            for c in range(4):
                ro, io, co = o.solve_residuals(Y[:, c].copy(), tol=tol, maxiter=10000)
                rt = o.solve_residuals(Y[:, c].copy(), tol=1e-13, maxiter=100000)[0]
                _say(f"  tol {tol:g} col {c}: iters {its[c]} (oracle {io}) err vs oracle {_rel(cols[c], ro):.2e} "
                     f"vs exact {_rel(cols[c], rt):.2e} (oracle {_rel(ro, rt):.2e})")
                assert cvs[c]
                if tol == 1e-12:
                    assert _rel(cols[c], ro) <= RES_RTOL
                else:
                    assert _rel(cols[c], rt) <= 3 * _rel(ro, rt) + 1e-12


def test_mixed_precision_coefficients(gpu):
    _say("case: mixed coefficients")
    refs, n, z, w, y = _rand_problem(62, 20_000, (2000, 150))
    o = oracle.OracleSolver(refs, n, z, w)
    co, io, cvo = o.solve_coefficients(y, tol=1e-12, maxiter=10000)
    with api.FixedEffectSolverROCm(_fes(refs, n, z), w, "mixed") as s:
        cg, ig, cvg = s.solve_coefficients(y, tol=1e-12, maxiter=10000)
    assert cvg and cvo
    assert _rel(sum(cg), sum(co)) <= 1e-8


def test_coefficients_match_oracle(gpu):
    _say("case: coefficients")
    refs, n, z, w, y = _rand_problem(61, 20_000, (2000, 150))
    o = oracle.OracleSolver(refs, n, z, w)
    co, io, cvo = o.solve_coefficients(y, tol=1e-10, maxiter=10000)
    with api.FixedEffectSolverROCm(_fes(refs, n, z), w, "fp64") as s:
        cg, ig, cvg = s.solve_coefficients(y, tol=1e-10, maxiter=10000)
    assert cvg and cvo and abs(ig - io) <= 1
    fit_o = sum(co)
    fit_g = sum(cg)
    assert _rel(fit_g, fit_o) <= 1e-8
    for a, b in zip(cg, co):
        np.testing.assert_allclose(a, b, rtol=1e-6, atol=1e-6)


def test_update_weights_matches_fresh_solver(gpu):
    # test/solve.jl:45-55 computes these but does not assert them; here they are asserted
    _say("case: update_weights")
    p1 = np.array(GOLD["p1"])
    x = np.array(GOLD["x"])
    fes = [api.FixedEffect(p1)]
    with api.FixedEffectSolverROCm(fes, np.ones(10), "fp64") as feM:
        for wv in (GOLD["weights_update_1"], GOLD["weights_update_2"]):
            wv = np.array(wv, np.float64)
            feM.update_weights(wv)
            r1, _, _ = feM.solve_residuals(x.copy())
            r2, _, _ = api.solve_residuals(x.copy(), fes, wv, double_precision=True)
            np.testing.assert_allclose(r1, r2, rtol=1e-14, atol=1e-15)
            ro, _, _ = oracle.solve_residuals(x, [p1.astype(np.int64)], [5], wv)
            np.testing.assert_allclose(r1, ro, rtol=1e-12, atol=1e-14)


def test_interaction_cases_converge(gpu):
    # test/solve.jl:57-66: interacted FE alone, and with the same refs uninteracted
    _say("case: interaction golden")
    p1 = np.array(GOLD["p1"])
    x = np.array(GOLD["x"])
    inter = np.array(GOLD["interaction"])
    for fes_def, zs in (([p1], [inter]), ([p1, p1], [None, inter])):
        fes = [api.FixedEffect(r, interaction=z) for r, z in zip(fes_def, zs)]
        r, it, cv = api.solve_residuals(x.copy(), fes, double_precision=True)
        assert cv
        ro, io, co = oracle.solve_residuals(x, [q.astype(np.int64) for q in fes_def], [5] * len(fes_def), None, zs)
        assert co
        np.testing.assert_allclose(r, ro, rtol=1e-8, atol=1e-10)


def test_maxiter_and_stop_codes(gpu):
    _say("case: maxiter")
    refs, n, z, w, y = _rand_problem(71, 10_000, (1000, 100))
    with api.FixedEffectSolverROCm(_fes(refs, n, z), w, "fp64") as s:
        r, it, cv = s.solve_residuals(y.copy(), tol=1e-14, maxiter=3)
        assert it == 3 and not cv
        r, it, cv = s.solve_residuals(y.copy(), tol=1e-14, maxiter=1)
        assert it == 1
        r, it, cv = s.solve_residuals(np.zeros(len(y)), tol=1e-8, maxiter=100)
        assert cv and it == 1 and not np.any(r)


def test_deterministic(gpu):
    # atomic-free fixed-order reductions: two solves are bitwise identical
    _say("case: determinism")
    p = synth.config("worker_firm_1e7", scale=0.01)
    with api.FixedEffectSolverROCm(_fes(p.refs, p.n), p.weights, "fp64") as s:
        a, ia, _ = s.solve_residuals(p.Y[:, 0].copy())
        b, ib, _ = s.solve_residuals(p.Y[:, 0].copy())
    assert ia == ib
    np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("k", [1, 5, 12, 30, 60])
def test_fixed_iteration_parity(gpu, k):
    # tol = 0: neither side stops early, so the k-th LSMR iterates and scalars
    # are compared directly (x0 = A'b as solve_residuals! uses).
    # * k <= 12: the same trajectory -- residuals and scalars to 1e-12.
    # * k = 30, 60: finite-precision LSMR is sensitive mid-run (Lanczos loses
    #   orthogonality and the timing of its repeated Ritz values depends on
    #   rounding; src/utils/lsmr.jl:113-192 is followed exactly).  The oracle
    #   run on permuted rows already departs from itself there (its cumulative
    #   normA estimate by ~0.2% at k = 30) and comes back later.  The bar is
    #   then equal progress: the GPU iterate's residual is at least as close to
    #   the converged residual as the oracle's (factor 3), and ||b - A x||,
    #   the objective, agrees to 1e-5.  The spreads are printed.
    _say(f"case: fixed iterations k={k}")
    p = synth.config("worker_firm_1e7", scale=0.02)
    y = p.Y[:, 0].copy()
    o = oracle.OracleSolver(p.refs, p.n, p.interactions, p.weights, nthreads=8)
    ro, io, _ = o.solve_residuals(y.copy(), tol=0.0, maxiter=k + 1)
    b = y * np.sqrt(p.weights)
    _, info = o.lsmr(o.mul_At(b), b, atol=0.0, btol=0.0, conlim=1e8, maxiter=k)
    perm = np.random.default_rng(5).permutation(p.N)
    op = oracle.OracleSolver([r[perm] for r in p.refs], p.n, [None] * 3, p.weights[perm], nthreads=8)
    rp = np.empty(p.N)
    rp[perm] = op.solve_residuals(y[perm].copy(), tol=0.0, maxiter=k + 1)[0]
    with api.FixedEffectSolverROCm(_fes(p.refs, p.n), p.weights, "fp64") as s:
        rg, ig, _ = s.solve_residuals(y.copy(), tol=0.0, maxiter=k + 1)
        gi = s.lsmr_info(0)
    _say(f"  iters {ig} {io} rel err {_rel(rg, ro):.3e} (the permuted-row oracle {_rel(rp, ro):.3e})")
    for key in ("normr", "normAr", "normA", "condA", "normx"):
        _say(f"  {key}: gpu {gi[key]!r} oracle {getattr(info, key)!r}")
    assert ig == io == k + 1
    if k <= 12:
        assert _rel(rg, ro) <= 1e-12
        for key in ("normr", "normAr", "normA", "normx"):
            assert abs(gi[key] - getattr(info, key)) <= 1e-12 * abs(getattr(info, key)), key
    else:
        rt = o.solve_residuals(y.copy(), tol=1e-13, maxiter=100_000)[0]
        _say(f"  distance to the converged residual: gpu {_rel(rg, rt):.3e} oracle {_rel(ro, rt):.3e} "
             f"permuted oracle {_rel(rp, rt):.3e}")
        assert _rel(rg, rt) <= 3 * _rel(ro, rt) + 1e-12
        assert abs(gi["normr"] - info.normr) <= 1e-5 * info.normr


def test_matvec_skewed(gpu):
    # Zipf-sized levels spanning many tiles (through-tile fix-up), empty levels, b != 0
    _say("case: matvec skewed")
    rng = np.random.default_rng(3)
    N = 300_000
    sizes = synth.zipf_sizes(N, 20_000, 1.1)
    r0 = np.repeat(np.arange(1, 20_001), sizes)
    rng.shuffle(r0)
    r1 = rng.integers(1, 3001, N)
    r1[r1 == 7] = 8
    r2 = rng.integers(1, 13, N)
    refs = [r0.astype(np.int64), r1.astype(np.int64), r2.astype(np.int64)]
    n = [20_000, 3000, 12]
    w = rng.lognormal(0, 0.5, N)
    o = oracle.OracleSolver(refs, n, None, w)
    x = rng.standard_normal(o.p)
    y = rng.standard_normal(N)
    with api.FixedEffectSolverROCm(_fes(refs, n), w, "fp64") as s:
        for a, b in ((1.0, 0.0), (0.3, -1.7)):
            np.testing.assert_allclose(s.mul_At(y, x, a, b), o.mul_At(y, x, a, b), rtol=1e-11, atol=1e-11)
            np.testing.assert_allclose(s.mul_A(x, y, a, b), o.mul_A(x, y, a, b), rtol=1e-12, atol=1e-12)


def test_schedule_and_autotune(gpu):
    # any schedule gives the same residuals up to the order of sums; the
    # autotuner returns a schedule no slower than the default it measured
    _say("case: schedules / autotune")
    p = synth.config("worker_firm_1e7", scale=0.02)
    fes = [api.FixedEffect.from_refs(r, k, z) for r, k, z in zip(p.refs, p.n, p.interactions)]
    y = np.ascontiguousarray(p.Y[:, 0])
    o = oracle.OracleSolver(p.refs, p.n, p.interactions, p.weights)
    ro = o.solve_residuals(y.copy(), tol=1e-12, maxiter=10000)[0]
    with api.FixedEffectSolverROCm(fes, p.weights, "fp64") as s:
        for sched in ((0, 0, -1, 0), (37, 1000, 0, 3), (300, 50, 1, 16)):
            s.set_schedule(*sched)
            r = s.solve_residuals(y.copy(), tol=1e-12, maxiter=10000)[0]
            _say(f"  schedule {sched}: rel err {_rel(r, ro):.2e}")
            assert _rel(r, ro) <= RES_RTOL
        best = s.autotune(K=1, reps=3)
        _say(f"  autotune: {best}")
        assert best["ms_per_iter"] <= best["default_ms_per_iter"]
        r = s.solve_residuals(y.copy(), tol=1e-12, maxiter=10000)[0]
        assert _rel(r, ro) <= RES_RTOL


def test_panels_match_single_solves(gpu):
    # uneven panels solved together give exactly the residuals and
    # iteration counts of solving each on its own handle
    _say("case: panels")
    from fixedeffects_rocm import device
    panels = [synth.uneven_panel(700 + i, 1e4, 8e4) for i in range(6)]
    solvers = [api.FixedEffectSolverROCm([api.FixedEffect.from_refs(r, k, z) for r, k, z in
                                          zip(p.refs, p.n, p.interactions)], p.weights, "fp64") for p in panels]
    try:
        single, bufs = [], []
        for p, s in zip(panels, solvers):
            b = device.DeviceBuffer(p.N * p.K)
            b.upload(np.asfortranarray(p.Y).ravel(order="F"))
            its, cvs = s.solve_residuals_device(b.address, p.K, p.N, tol=1e-8, maxiter=10000)
            single.append((b.download(p.N * p.K), its))
            b.upload(np.asfortranarray(p.Y).ravel(order="F"))
            bufs.append(b)
        its, cvs = api.solve_residuals_panels(solvers, [b.address for b in bufs], [p.K for p in panels],
                                              [p.N for p in panels], tol=1e-8, maxiter=10000)
        for i, (p, b) in enumerate(zip(panels, bufs)):
            r = b.download(p.N * p.K)
            _say(f"  panel {i}: N={p.N} K={p.K} iters {its[i]}")
            assert its[i] == single[i][1]
            assert all(cvs[i])
            np.testing.assert_array_equal(r, single[i][0])
    finally:
        for s in solvers:
            s.close()


def test_solve_more_than_32_columns(gpu):
    # 40 columns run as a batch of 32 and a batch of 8; every column matches
    # the oracle at a tight tolerance
    _say("case: 40 columns")
    p = synth.config("worker_firm_1e7_k32", scale=0.005)
    rng = np.random.default_rng(3)
    Y = np.concatenate([p.Y, p.Y[:, :8] + rng.standard_normal((p.N, 8))], axis=1)
    worst, ig, io = _check_solve(p.refs, p.n, p.interactions, p.weights, Y, tol=1e-12, iter_slack=10)
    _say(f"  worst rel err {worst:.2e}")
    assert worst <= RES_RTOL


def test_panels_of_mixed_precisions(gpu):
    # panels whose handles run FP64, FP32 and mixed precision in one call give
    # the residuals of solving each alone
    _say("case: panels of several precisions")
    from fixedeffects_rocm import device
    panels = [synth.uneven_panel(900 + i, 1e4, 4e4) for i in range(3)]
    precs = ["fp64", "fp32", "mixed"]
    solvers = [api.FixedEffectSolverROCm([api.FixedEffect.from_refs(r, k, z) for r, k, z in
                                          zip(p.refs, p.n, p.interactions)], p.weights, pr)
               for p, pr in zip(panels, precs)]
    try:
        single, bufs = [], []
        for p, s in zip(panels, solvers):
            b = device.DeviceBuffer(p.N * p.K)
            b.upload(np.asfortranarray(p.Y).ravel(order="F"))
            s.solve_residuals_device(b.address, p.K, p.N, tol=1e-8, maxiter=10000)
            single.append(b.download(p.N * p.K))
            b.upload(np.asfortranarray(p.Y).ravel(order="F"))
            bufs.append(b)
        its, cvs = api.solve_residuals_panels(solvers, [b.address for b in bufs], [p.K for p in panels],
                                              [p.N for p in panels], tol=1e-8, maxiter=10000)
        for i, (p, b) in enumerate(zip(panels, bufs)):
            _say(f"  panel {i} ({precs[i]}): K={p.K} iters {its[i]}")
            np.testing.assert_array_equal(b.download(p.N * p.K), single[i])
    finally:
        for s in solvers:
            s.close()


@pytest.mark.parametrize("prec", ["fp64", "mixed", "fp32"])
def test_wide_batch_few_level_fes(gpu, prec):
    # K = 16 with a 20-level FE and a 45-level interacted FE: in FP64 the
    # fused pass's per-column phased accumulators and the set-up pass on the
    # same partition; in FP32 (65 lane-accumulator slots exceed 64 KB) the
    # separate few-level pass
    _say(f"case: K=16 few-level FEs {prec}")
    rng = np.random.default_rng(77)
    refs, n, z, w, y = _rand_problem(78, 30_000, (3000, 20, 45), interaction=(2,))
    K = 16
    Y = np.column_stack([y] + [rng.standard_normal(30_000) for _ in range(K - 1)])
    o = oracle.OracleSolver(refs, n, z, w, nthreads=8)
    tight = [o.solve_residuals(Y[:, c].copy(), tol=1e-13, maxiter=100_000)[0] for c in range(K)]
    tol, bar = (1e-6, 1e-3) if prec == "fp32" else (1e-12, RES_RTOL)
    with api.FixedEffectSolverROCm(_fes(refs, n, z), w, prec) as s:
        cols = [Y[:, c].copy() for c in range(K)]
        _, its, cvs = s.solve_residuals_many(cols, tol=tol, maxiter=100_000)
    worst = max(_rel(a, b) for a, b in zip(cols, tight))
    _say("  iters", its[:4], "worst", worst)
    assert all(cvs) and worst <= bar


@pytest.mark.parametrize("algorithm,K", [("lsmr", 1), ("lsmr", 4), ("lsmr", 16), ("projected", 1)])
def test_csr_levels_longer_than_tiles(gpu, algorithm, K):
    # a CSR FE with Zipf-sized levels, the largest spanning many CSR tiles:
    # CSR tiles are cut at segment heads except inside such levels, whose
    # partials the fix-up launch after the CSR pass combines (here also with
    # small CSR tiles, which cut many more levels); the converged residuals
    # match the oracle's tight solve
    _say(f"case: CSR levels longer than tiles {algorithm} K={K}")
    rng = np.random.default_rng(40 + K)
    N = 200_000 if K < 16 else 100_000
    r0 = rng.integers(1, 20_001, N)  # most levels: FE s
    sizes = synth.zipf_sizes(N, 3000, 1.1)
    r1 = np.repeat(np.arange(1, 3001), sizes)
    rng.shuffle(r1)
    r2 = rng.integers(1, 13, N)
    refs = [r0.astype(np.int64), r1.astype(np.int64), r2.astype(np.int64)]
    n = [20_000, 3000, 12]
    w = rng.lognormal(0, 0.5, N)
    Y = rng.standard_normal((N, K))
    o = oracle.OracleSolver(refs, n, None, w, nthreads=8)
    tight = [o.solve_residuals(Y[:, c].copy(), tol=1e-13, maxiter=100_000)[0] for c in range(K)]
    with api.FixedEffectSolverROCm(_fes(refs, n), w, "fp64", algorithm=algorithm) as s:
        for sched in ((0, 0, -1, 0), (0, 3000, -1, 0)):
            s.set_schedule(*sched)
            cols = [np.ascontiguousarray(Y[:, c]).copy() for c in range(K)]
            _, its, cvs = s.solve_residuals_many(cols, tol=1e-12, maxiter=100_000)
            worst = max(_rel(cols[c], tight[c]) for c in range(K))
            _say(f"  schedule {sched}: iters {its[:4]} worst rel err {worst:.2e}")
            assert all(cvs) and worst <= RES_RTOL


@pytest.mark.parametrize("algorithm,K", [("lsmr", 1), ("lsmr", 16), ("projected", 1)])
def test_separate_csr_and_fixup_launches(gpu, monkeypatch, algorithm, K):
    # FEHIP_SEG_FIX=0 (read when a handle is created): the CSR pass and the
    # fix-up as separate launches (at K >= 16 the fix-up's FE s part on a
    # second graph branch) give the merged launch's converged residuals
    _say(f"case: separate CSR / fix-up launches {algorithm} K={K}")
    p = synth.config("worker_firm_1e7", scale=0.01)
    rng = np.random.default_rng(7)
    Y = np.column_stack([p.Y[:, 0]] + [rng.standard_normal(p.N) for _ in range(K - 1)])
    o = oracle.OracleSolver(p.refs, p.n, p.interactions, p.weights, nthreads=8)
    tight = [o.solve_residuals(Y[:, c].copy(), tol=1e-13, maxiter=100_000)[0] for c in range(K)]
    out = {}
    for sf in ("1", "0"):
        monkeypatch.setenv("FEHIP_SEG_FIX", sf)
        with api.FixedEffectSolverROCm(_fes(p.refs, p.n, p.interactions), p.weights, "fp64",
                                       algorithm=algorithm) as s:
            cols = [np.ascontiguousarray(Y[:, c]).copy() for c in range(K)]
            _, its, cvs = s.solve_residuals_many(cols, tol=1e-12, maxiter=100_000)
        worst = max(_rel(cols[c], tight[c]) for c in range(K))
        _say(f"  FEHIP_SEG_FIX={sf}: iters {its[:4]} worst rel err {worst:.2e}")
        assert all(cvs) and worst <= RES_RTOL
        out[sf] = cols
    for a, b in zip(out["1"], out["0"]):
        assert _rel(a, b) <= 1e-9


@pytest.mark.parametrize("prec,K", [("fp32", 1), ("fp32", 16), ("mixed", 16)])
def test_fp32_lane_accumulators_in_fused_pass(gpu, prec, K):
    # one 50-level interacted FE beside FE s and a CSR FE: in FP32 its plain
    # float lane accumulators (50 slots, 51 KB per workgroup) fit the fused
    # pass; converged residuals within the precision's bar of the exact ones
    _say(f"case: FP32 lane accumulators {prec} K={K}")
    rng = np.random.default_rng(90 + K)
    refs, n, z, w, y = _rand_problem(91, 40_000, (4000, 300, 50), interaction=(2,))
    Y = np.column_stack([y] + [rng.standard_normal(40_000) for _ in range(K - 1)])
    o = oracle.OracleSolver(refs, n, z, w, nthreads=8)
    tight = [o.solve_residuals(Y[:, c].copy(), tol=1e-13, maxiter=100_000)[0] for c in range(K)]
    tol, bar = (1e-6, 1e-3) if prec == "fp32" else (1e-12, RES_RTOL)
    with api.FixedEffectSolverROCm(_fes(refs, n, z), w, prec) as s:
        cols = [Y[:, c].copy() for c in range(K)]
        _, its, cvs = s.solve_residuals_many(cols, tol=tol, maxiter=100_000)
    worst = max(_rel(a, b) for a, b in zip(cols, tight))
    _say("  iters", its[:4], "worst", worst)
    assert all(cvs) and worst <= bar
