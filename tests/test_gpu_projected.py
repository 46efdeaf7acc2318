"""GPU tests of the projected LSMR (FEHIP_LSMR_PROJECTED: the FE of most
levels eliminated, Frisch-Waugh-Lovell) against the oracle.

The projected solve minimizes the same least-squares objective, so its
residual is the reference's residual: converged (tol 1e-12) it must agree with
the oracle's converged residual to 1e-8 relative, the bar of the plain path.
At the reference tol 1e-8 it must be at least as close to the exact residual
(an oracle solve at tol 1e-13) as the oracle's own tol-1e-8 residual, within
the factor 3 the plain path is held to.  Its iteration counts are not the
reference's (that is the point: fewer), so they are only checked to be lower.
"""
import sys

import numpy as np
import pytest

import oracle
from fixedeffects_rocm import api, synth

pytestmark = pytest.mark.gpu

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
    for k in n:
        r = rng.integers(1, k + 1, N).astype(np.int64)
        if N >= k:
            r[:k] = np.arange(1, k + 1)
            rng.shuffle(r)
        if empty_level and k > 2:
            r[r == 2] = 1
        refs.append(r)
    z = [rng.uniform(0.5, 1.5, N) if g in interaction else None for g in range(len(n))]
    w = rng.lognormal(0, 0.5, N) if weights else None
    y = rng.standard_normal((N, 1))
    return refs, list(n), z, w, y


def _solve(refs, n, z, w, Y, algorithm, prec="fp64", tol=1e-8):
    with api.FixedEffectSolverROCm(_fes(refs, n, z), w, prec, algorithm=algorithm) as s:
        cols = [np.ascontiguousarray(Y[:, c]).copy() for c in range(Y.shape[1])]
        _, its, cvs = s.solve_residuals_many(cols, tol=tol, maxiter=100_000)
    return cols, its, cvs


def _check(refs, n, z, w, Y, prec="fp64", tol=1e-8, fewer=True):
    """Projected vs the oracle: at `tol` as accurate as the oracle's own solve
    (against a tol-1e-13 oracle solve); converged (1e-12) within 1e-8."""
    o = oracle.OracleSolver(refs, n, z, w, nthreads=8)
    K = Y.shape[1]
    tight = [o.solve_residuals(Y[:, c].copy(), tol=1e-13, maxiter=100_000)[0] for c in range(K)]
    ref = [o.solve_residuals(Y[:, c].copy(), tol=tol, maxiter=100_000) for c in range(K)]
    rp, ip, cp = _solve(refs, n, z, w, Y, "projected", prec, tol)
    _, il, _ = _solve(refs, n, z, w, Y, "lsmr", prec, tol)
    for c in range(K):
        assert cp[c]
        assert _rel(rp[c], tight[c]) <= 3 * _rel(ref[c][0], tight[c]) + 1e-12, (c, _rel(rp[c], tight[c]),
                                                                                 _rel(ref[c][0], tight[c]))
        if fewer:
            assert ip[c] < il[c], (c, ip[c], il[c])
    rc, _, cc = _solve(refs, n, z, w, Y, "projected", prec, 1e-12)
    worst = max(_rel(rc[c], tight[c]) for c in range(K))
    assert all(cc) and worst <= RES_RTOL, worst
    return ip, il, worst


def test_projected_worker_firm(gpu):
    _say("case: projected worker_firm scale 0.02")
    p = synth.config("worker_firm_1e7", scale=0.02)
    ip, il, worst = _check(p.refs, p.n, p.interactions, p.weights, p.Y)
    _say("  iters projected", ip, "plain", il, "converged rel err", worst)
    assert ip[0] <= 0.8 * il[0]


def test_projected_two_way(gpu):
    _say("case: projected cpu_two_way")
    p = synth.config("cpu_two_way")
    ip, il, worst = _check(p.refs, p.n, p.interactions, p.weights, p.Y)
    _say("  iters projected", ip, "plain", il, "converged rel err", worst)


@pytest.mark.parametrize("inter", [(2,), (0,), (0, 1, 2)])
def test_projected_interactions(gpu, inter):
    # interactions switch off the factored form; (0,) interacts FE s itself
    _say(f"case: projected interactions {inter}")
    refs, n, z, w, y = _rand_problem(71, 30_000, (3000, 200, 9), interaction=inter)
    ip, il, worst = _check(refs, n, z, w, y, fewer=False)
    _say("  iters projected", ip, "plain", il, "converged rel err", worst)


def test_projected_empty_levels_and_unit_weights(gpu):
    _say("case: projected empty levels, unit weights")
    refs, n, z, w, y = _rand_problem(72, 20_000, (2500, 300), weights=False, empty_level=True)
    _check(refs, n, z, w, y, fewer=False)


def test_projected_batched(gpu):
    _say("case: projected K=32")
    p = synth.config("worker_firm_1e7_k32", scale=0.01)
    ip, il, worst = _check(p.refs, p.n, p.interactions, p.weights, p.Y[:, :32])
    _say("  converged rel err", worst, "mean iters", np.mean(ip), np.mean(il))


def test_projected_mixed_and_fp32(gpu):
    _say("case: projected mixed / fp32")
    p = synth.config("worker_firm_1e7", scale=0.02)
    o = oracle.OracleSolver(p.refs, p.n, p.interactions, p.weights, nthreads=8)
    y = p.Y[:, :1]
    tight = o.solve_residuals(y[:, 0].copy(), tol=1e-13, maxiter=100_000)[0]
    rm, im, cm = _solve(p.refs, p.n, p.interactions, p.weights, y, "projected", "mixed", 1e-12)
    assert cm[0] and _rel(rm[0], tight) <= RES_RTOL
    ro32, io32, _ = oracle.OracleSolver(p.refs, p.n, p.interactions, p.weights, double_precision=False,
                                        nthreads=8).solve_residuals(y[:, 0].copy(), tol=1e-6, maxiter=100_000)
    r32, i32, c32 = _solve(p.refs, p.n, p.interactions, p.weights, y, "projected", "fp32", 1e-6)
    assert c32[0]
    assert _rel(r32[0], tight) <= max(3 * _rel(ro32, tight), 1e-4)
    _say("  mixed iters", im, "fp32 iters", i32, "oracle fp32", io32)


def test_projected_panels_and_x0(gpu):
    # panels of projected handles give the single-handle results bitwise; a
    # start from given coefficients reaches the same converged residual
    _say("case: projected panels / x0")
    from fixedeffects_rocm import device, warmstart
    panels = [synth.uneven_panel(900 + i, 1e4, 6e4) for i in range(4)]
    solvers = [api.FixedEffectSolverROCm([api.FixedEffect.from_refs(r, k, z) for r, k, z in
                                          zip(p.refs, p.n, p.interactions)], p.weights, "fp64", algorithm="projected")
               for p in panels]
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
            assert its[i] == single[i][1] and all(cvs[i])
            np.testing.assert_array_equal(b.download(p.N * p.K), single[i][0])
        p, s = panels[0], solvers[0]
        o = oracle.OracleSolver(p.refs, p.n, p.interactions, p.weights, nthreads=8)
        y = np.ascontiguousarray(p.Y[:, 0])
        tight = o.solve_residuals(y.copy(), tol=1e-13, maxiter=100_000)[0]
        col = y.copy()
        x0 = np.random.default_rng(5).standard_normal(int(sum(p.n)))
        warmstart.solve_residuals_from(s, [col], [x0], tol=1e-12, maxiter=100_000)
        assert _rel(col, tight) <= RES_RTOL
    finally:
        for s in solvers:
            s.close()


@pytest.mark.parametrize("prec,K,inter", [("fp64", 8, (2,)), ("fp64", 16, ()), ("fp32", 4, ())])
def test_projected_batches_of_other_forms(gpu, prec, K, inter):
    # batched projected solves in the non-factored form (interaction), at the
    # width where the fix-up splits (K = 16) and in FP32: converged, every
    # column within the precision's bar of the exact residual
    _say(f"case: projected {prec} K={K} interactions {inter}")
    rng = np.random.default_rng(100 + K)
    refs, n, z, w, y = _rand_problem(90 + K, 25_000, (2500, 250, 12), interaction=inter)
    Y = np.column_stack([y[:, 0]] + [rng.standard_normal(25_000) for _ in range(K - 1)])
    o = oracle.OracleSolver(refs, n, z, w, nthreads=8)
    tight = [o.solve_residuals(Y[:, c].copy(), tol=1e-13, maxiter=100_000)[0] for c in range(K)]
    tol, bar = (1e-12, RES_RTOL) if prec == "fp64" else (1e-6, 1e-3)
    rp, ip, cp = _solve(refs, n, z, w, Y, "projected", prec, tol)
    assert all(cp)
    worst = max(_rel(rp[c], tight[c]) for c in range(K))
    _say("  iters", ip[:4], "worst rel err", worst)
    assert worst <= bar


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_projected_coefficients(gpu, prec):
    # solve_coefficients! on the projected problem: FE s's coefficients are
    # recovered as the least-squares fit of the rest; after normalize! the
    # coefficients and the fitted values match the oracle's reference solve
    _say(f"case: projected coefficients {prec}")
    refs, n, z, w, y = _rand_problem(61, 20_000, (2000, 150))
    y = y[:, 0]
    o = oracle.OracleSolver(refs, n, z, w)
    tol = 1e-10 if prec == "fp64" else 1e-6
    co, io, cvo = o.solve_coefficients(y.copy(), tol=tol, maxiter=10000)
    with api.FixedEffectSolverROCm(_fes(refs, n, z), w, prec, algorithm="projected") as s:
        cg, ig, cvg = s.solve_coefficients(y.copy(), tol=tol, maxiter=10000)
        s.set_algorithm("lsmr")
        _, il, _ = s.solve_coefficients(y.copy(), tol=tol, maxiter=10000)
    _say("  iters projected", ig, "plain", il, "oracle", io)
    assert cvg and cvo and ig < il
    bar = 1e-8 if prec == "fp64" else 1e-4
    assert _rel(sum(cg), sum(co)) <= bar
    for a, b in zip(cg, co):
        np.testing.assert_allclose(a, b, rtol=1e-6 if prec == "fp64" else 1e-3, atol=1e-6 if prec == "fp64" else 1e-3)
