"""GPU checks of the learned warm start: the bf16-MFMA network against its FP64
restatement (forward, gradients, training), the device-computed features
against a host recomputation, and LSMR from given / learned starts."""
import sys

import numpy as np
import pytest

import oracle
from oracle import mlp_oracle as M
from fixedeffects_rocm import api, synth, warmstart

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _params(seed, scale=0.15):
    rng = np.random.default_rng(seed)
    _, n = M.layout()
    p = rng.standard_normal(n) * scale
    Ws, bs = M.unpack(p)
    Ws[3][:, 1:] = 0.0
    bs[3][1:] = 0.0
    return M.pack(Ws, bs).astype(np.float32)


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, float) - np.asarray(b, float)) / max(np.linalg.norm(b), 1e-300))


def test_mlp_forward_matches_fp64(gpu):
    # also pins the MFMA operand layout: a wrong lane mapping gives O(1) errors
    _say("case: mlp forward")
    net = warmstart.WarmStartNet(seed=3)
    p = _params(1)
    net.params = p
    rng = np.random.default_rng(2)
    for n in (16, 1000, 70_000):
        X = rng.standard_normal((n, 16)).astype(np.float32)
        y = net.predict(X)
        ye = M.forward(p.astype(np.float64), X, emulate_bf16=True)[0]
        yx = M.forward(p.astype(np.float64), X, emulate_bf16=False)[0]
        _say(f"  n={n} rel err vs bf16-emulating oracle {_rel(y, ye):.2e}, vs exact FP64 {_rel(y, yx):.2e}")
        assert _rel(y, ye) <= 2e-3
        assert _rel(y, yx) <= 3e-2


def test_mlp_gradient_matches_fp64(gpu):
    _say("case: mlp gradient")
    net = warmstart.WarmStartNet(seed=4)
    p = _params(5)
    net.params = p
    rng = np.random.default_rng(6)
    X = rng.standard_normal((3000, 16)).astype(np.float32)
    t = rng.standard_normal(3000).astype(np.float32)
    loss, g = net.gradient(X, t)
    le, ge = M.loss_and_grad(p.astype(np.float64), X, t, emulate_bf16=True)
    lx, gx = M.loss_and_grad(p.astype(np.float64), X, t, emulate_bf16=False)
    offs, _ = M.layout()
    for l, ((ow, ob), dout) in enumerate(zip(offs, M.D_OUT)):
        for name, sl in (("W", slice(ow, ob)), ("b", slice(ob, ob + dout))):
            e_em = _rel(g[sl], ge[sl])
            e_ex = _rel(g[sl], gx[sl])
            _say(f"  layer {l} {name}: rel err vs emulating oracle {e_em:.2e}, vs exact {e_ex:.2e}")
            assert e_em <= 5e-3
            assert e_ex <= 6e-2
    assert abs(loss - le) <= 1e-3 * le


def test_mlp_fit_learns(gpu):
    _say("case: mlp fit")
    rng = np.random.default_rng(7)
    X = rng.standard_normal((20_000, 16)).astype(np.float32)
    t = (np.tanh(X[:, 0]) + 0.5 * X[:, 1] * X[:, 2] - 0.3 * X[:, 3]).astype(np.float32)
    net = warmstart.WarmStartNet(seed=8)
    l0 = net.gradient(X, t)[0]
    losses = net.fit(X, t, epochs=8, batch=1024, lr=2e-3, seed=9)
    _say(f"  loss {l0:.4f} -> {losses[-1]:.4f} ({losses})")
    assert losses[-1] < 0.15 * l0


def _host_features(p, y):
    w = p.weights if p.weights is not None else np.ones(len(y))
    b = np.sqrt(w) * y
    sig = np.sqrt(np.mean(b * b))
    out = []
    r = y.copy()
    for g, (ref, k) in enumerate(zip(p.refs, p.n)):
        idx = ref - 1
        cnt = np.bincount(idx, minlength=k)
        sw = np.bincount(idx, weights=w, minlength=k)
        mean = np.divide(np.bincount(idx, weights=w * y, minlength=k), sw, out=np.zeros(k), where=sw > 0)
        apg = np.divide(np.bincount(idx, weights=w * r, minlength=k), sw, out=np.zeros(k), where=sw > 0)
        r = r - apg[idx]
        out.append((cnt, mean / sig, apg / sig))
    return out, sig


def test_features_match_host(gpu):
    _say("case: features")
    p = synth.uneven_panel(123, 2e4, 5e4)
    y = p.Y[:, 0].copy()
    fes = [api.FixedEffect.from_refs(r, k) for r, k in zip(p.refs, p.n)]
    with api.FixedEffectSolverROCm(fes, p.weights, "fp64") as s:
        F, ap, sig = warmstart.features(s, y)
    host, hsig = _host_features(p, y)
    assert abs(sig - hsig) <= 1e-12 * hsig
    off = 0
    for g, (cnt, mean, apg) in enumerate(host):
        k = len(cnt)
        blk = F[off:off + k]
        np.testing.assert_allclose(blk[:, 0], np.log1p(cnt), rtol=1e-6)
        np.testing.assert_allclose(blk[:, 1], mean, rtol=2e-5, atol=2e-6)
        np.testing.assert_allclose(blk[:, 3], apg, rtol=2e-5, atol=2e-6)
        np.testing.assert_allclose(ap[off:off + k], apg * sig, rtol=1e-10, atol=1e-12)
        assert np.all(blk[:, 6 + g] == 1.0)
        off += k
    # degree of the first FE: distinct second-FE levels per level
    pairs = np.unique(np.stack([p.refs[0], p.refs[1]], 1), axis=0)
    deg = np.bincount(pairs[:, 0] - 1, minlength=p.n[0])
    np.testing.assert_allclose(F[:p.n[0], 2], np.log1p(deg), rtol=1e-6)


def test_start_from_coefficients(gpu):
    # LSMR from the exact solution stops at once; from any start the
    # residuals agree.  (One alternating-projection sweep does not by itself
    # save Krylov iterations: it damps the error components LSMR removes
    # quickly anyway, so its count is only reported.)
    _say("case: start from coefficients")
    p = synth.uneven_panel(321, 5e4, 1e5)
    y = p.Y[:, 0].copy()
    fes = [api.FixedEffect.from_refs(r, k) for r, k in zip(p.refs, p.n)]
    with api.FixedEffectSolverROCm(fes, p.weights, "fp64") as s:
        F, ap, sig = warmstart.features(s, y)
        exact, _, _ = warmstart.solve_coefficients_from(s, y, ap, tol=1e-13)
        exact = np.concatenate(exact)
        r_ref = y.copy()
        it_ref, _ = warmstart.solve_residuals_from(s, [r_ref], None, tol=1e-10)
        r_ap = y.copy()
        it_ap, _ = warmstart.solve_residuals_from(s, [r_ap], [ap], tol=1e-10)
        r_ex = y.copy()
        it_ex, _ = warmstart.solve_residuals_from(s, [r_ex], [exact], tol=1e-10)
    _say(f"  iterations: reference start {it_ref}, alternating projection {it_ap}, exact {it_ex}")
    assert it_ex[0] <= 3
    assert _rel(r_ap, r_ref) <= 1e-8 and _rel(r_ex, r_ref) <= 1e-8


@pytest.mark.parametrize("seed", [900001, 900002])
def test_learned_start_residuals_agree(gpu, seed):
    # the network's start changes only where LSMR begins: the residuals
    # agree with the reference start's, for one column (whose feature pass
    # shares the solve's buffers) and for several
    _say(f"case: learned start, panel {seed}")
    p = synth.uneven_panel(seed, 1e4, 6e4)
    fes = [api.FixedEffect.from_refs(r, k, z) for r, k, z in zip(p.refs, p.n, p.interactions)]
    net = warmstart.WarmStartNet(seed=11)
    with api.FixedEffectSolverROCm(fes, p.weights, "fp64") as s:
        ref = [np.ascontiguousarray(p.Y[:, c]).copy() for c in range(p.K)]
        it_ref, cv_ref = warmstart.solve_residuals_from(s, ref, None, tol=1e-11)
        lrn = [np.ascontiguousarray(p.Y[:, c]).copy() for c in range(p.K)]
        it_l, cv_l = warmstart.solve_residuals_learned(s, lrn, net, tol=1e-11)
    errs = [_rel(a, b) for a, b in zip(lrn, ref)]
    _say(f"  K={p.K} iterations reference {it_ref} learned {it_l} rel diff {max(errs):.2e}")
    assert all(cv_ref) and all(cv_l)
    assert max(errs) <= 1e-7
