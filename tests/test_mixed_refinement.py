"""Why the mixed mode's second phase is FP64 LSMR rather than iterative
refinement with FP32 corrections (CPU emulation, no GPU).

Refinement would repeat: r = b - A x in FP64, round r to FP32 for the FP32
streams, solve min ||r - A d|| with FP32-stream LSMR to a tolerance relative
to the current error, x += d.  For the least-squares problems of
residualization (b far from range(A): the residual is the result) every
correction solve works on an r whose size is ||r*||, and the FP32 streams'
rounding, of order eps32 * ||r||, keeps the normal-equation test that LSMR
stops on (reference src/utils/lsmr.jl:198-215: ||A'r|| / (||A|| ||r||)) near
1e-8 -- the tolerance itself -- however many corrections follow.  The mixed
mode instead runs one FP32 phase to 1e-6 and finishes in FP64, which has no
such floor.

The emulation follows the GPU's FP32 phase: matrix entries and every vector
(u, v, h, hbar, x) stored as float32, matrix-vector sums of exact products
in float64, scalars in float64.  LSMR is the textbook algorithm (Fong and
Saunders, SIAM J. Sci. Comput. 33(5), 2011, algorithm 1 with damp = 0 and
its ||r|| estimate, section 4).  The panel is the scaled worker/firm/year
workload, columns scaled to unit norm as the reference's scale! does
(src/FixedEffect.jl:106-140).
"""
import numpy as np
import scipy.sparse as sp

from fixedeffects_rocm import synth


def _f32(a):
    return a.astype(np.float32).astype(np.float64)


def _problem(scale, seed=1):
    P = synth.config("worker_firm_1e7", scale=scale, seed=seed)
    N = P.N
    sw = np.sqrt(P.weights)
    cols = []
    for r, n in zip(P.refs, P.n):
        M = sp.csr_matrix((sw, (np.arange(N), r - 1)), shape=(N, n))
        nrm = np.sqrt(np.asarray(M.multiply(M).sum(axis=0)).ravel())
        cols.append(M @ sp.diags(1.0 / nrm))
    return sp.hstack(cols).tocsr(), sw * P.Y[:, 0]


def _lsmr(A, b, tol, low, x0=None, maxit=10_000):
    """Returns (x, iterations, ||A|| estimate).  low: FP32 storage."""
    rd = _f32 if low else (lambda a: a)
    if low:
        A = sp.csr_matrix((_f32(A.data), A.indices, A.indptr), shape=A.shape)
    At = A.T.tocsr()
    x = np.zeros(A.shape[1]) if x0 is None else x0.copy()
    normb = np.linalg.norm(b)
    u = rd(b - A @ x)
    beta = np.linalg.norm(u)
    u = rd(u / beta)
    v = rd(At @ u)
    alpha = np.linalg.norm(v)
    v = rd(v / alpha)
    zetabar, alphabar, rho, rhobar, cbar, sbar = alpha * beta, alpha, 1.0, 1.0, 1.0, 0.0
    h, hbar, x = v.copy(), np.zeros_like(v), rd(x)
    betadd, betad, rhodold, tautildeold, thetatilde, zeta = beta, 0.0, 1.0, 0.0, 0.0, 0.0
    normA2 = alpha * alpha
    for it in range(1, maxit + 1):
        u = rd(rd(A @ v) - alpha * u)
        beta = np.linalg.norm(u)
        u = rd(u / beta)
        v = rd(rd(At @ u) - beta * v)
        alpha = np.linalg.norm(v)
        v = rd(v / alpha)
        rhoold, rho = rho, np.hypot(alphabar, beta)
        c, s = alphabar / rho, beta / rho
        thetanew, alphabar = s * alpha, c * alpha
        rhobarold, zetaold, thetabar = rhobar, zeta, sbar * rho
        rhotemp = cbar * rho
        rhobar = np.hypot(rhotemp, thetanew)
        cbar, sbar = rhotemp / rhobar, thetanew / rhobar
        zeta, zetabar = cbar * zetabar, -sbar * zetabar
        hbar = rd(h - (thetabar * rho / (rhoold * rhobarold)) * hbar)
        x = rd(x + (zeta / (rho * rhobar)) * hbar)
        h = rd(v - (thetanew / rho) * h)
        betahat, betadd = c * betadd, -s * betadd
        thetatildeold = thetatilde
        rhotildeold = np.hypot(rhodold, thetabar)
        ctildeold, stildeold = rhodold / rhotildeold, thetabar / rhotildeold
        thetatilde, rhodold = stildeold * rhobar, ctildeold * rhobar
        betad = -stildeold * betad + ctildeold * betahat
        tautildeold = (zetaold - thetatildeold * tautildeold) / rhotildeold
        taud = (zeta - thetatilde * tautildeold) / rhodold
        normr = np.sqrt((betad - taud) ** 2 + betadd ** 2)
        normA2 += beta * beta
        normA = np.sqrt(normA2)
        normA2 += alpha * alpha
        test2 = abs(zetabar) / (normA * normr)
        if test2 <= tol or normr / normb <= tol + tol * normA * np.linalg.norm(x) / normb:
            break
    return x, it, normA


def _test2(A, b, x, normA):
    r = b - A @ x
    return np.linalg.norm(A.T @ r) / (normA * np.linalg.norm(r))


def test_fp32_corrections_stall_at_the_tolerance():
    A, b = _problem(0.005)  # N = 50,000, 5,520 levels
    x64, n64, normA = _lsmr(A, b, 1e-8, False)
    assert _test2(A, b, x64, normA) <= 1.5e-8
    # the mixed mode: FP32 phase to 1e-6, FP64 phase from its x
    x1, n1, _ = _lsmr(A, b, 1e-6, True)
    x2, n2, _ = _lsmr(A, b, 1e-8, False, x0=x1)
    assert _test2(A, b, x2, normA) <= 1.5e-8
    # refinement: each correction to 1% of the current test value
    x = np.zeros(A.shape[1])
    hist, n32 = [], 0
    for _ in range(6):
        t = _test2(A, b, x, normA)
        hist.append(t)
        d, n, _ = _lsmr(A, b - A @ x, max(0.01 * t, 1e-7), True)
        x, n32 = x + d, n32 + n
    hist.append(_test2(A, b, x, normA))
    print(f"FP64 {n64} it; mixed {n1} FP32 + {n2} FP64 it; refinement {n32} FP32 it, test2 "
          + " ".join(f"{h:.1e}" for h in hist))
    # after six corrections (about as many FP32 iterations as the mixed
    # mode's two phases together) the test is still above the tolerance,
    # and the last corrections barely move it
    assert n32 >= 0.9 * (n1 + n2)
    assert hist[-1] > 1e-8, hist
    assert hist[-1] > 0.3 * hist[-3], hist
