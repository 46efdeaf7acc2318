"""Pins the FP64 network restatement (oracle/mlp_oracle.py) by central finite
differences, so it can check the GPU network's gradients."""
import numpy as np

from oracle import mlp_oracle as M


def _params(seed):
    rng = np.random.default_rng(seed)
    _, n = M.layout()
    p = rng.standard_normal(n) * 0.2
    Ws, bs = M.unpack(p)
    Ws[3][:, 1:] = 0.0  # only the first output column is live
    bs[3][1:] = 0.0
    return M.pack(Ws, bs)


def test_pack_roundtrip():
    p = _params(0)
    Ws, bs = M.unpack(p)
    np.testing.assert_array_equal(M.pack(Ws, bs), p)
    assert [W.shape for W in Ws] == [(16, 128), (128, 128), (128, 128), (128, 16)]


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = _params(2)
    X = rng.standard_normal((40, 16))
    t = rng.standard_normal(40)
    _, g = M.loss_and_grad(p, X, t)
    offs, n = M.layout()
    # a sample of entries from every weight block and bias
    picks = []
    for (ow, ob), din, dout in zip(offs, M.D_IN, M.D_OUT):
        picks += [ow + int(rng.integers(0, din * dout)) for _ in range(6)]
        picks += [ob + int(rng.integers(0, min(dout, 16 if dout == 16 else dout)))]
    picks = [q for q in picks if not (offs[3][0] <= q < offs[3][1] and (q - offs[3][0]) % 16 != 0)]
    eps = 1e-6
    for q in picks:
        pp, pm = p.copy(), p.copy()
        pp[q] += eps
        pm[q] -= eps
        fd = (M.loss_and_grad(pp, X, t)[0] - M.loss_and_grad(pm, X, t)[0]) / (2 * eps)
        assert abs(fd - g[q]) <= 1e-6 * max(1.0, abs(fd)), (q, fd, g[q])


def test_bf16_rounding():
    x = np.array([1.0, 1.00390625, 1.005859375, -3.1415926, 0.0])
    r = M.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0  # tie to even
    assert r[2] == 1.0078125
    assert abs(r[3] - -3.140625) < 1e-12
    assert r[4] == 0.0
