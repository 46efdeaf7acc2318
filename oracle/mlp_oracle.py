"""FP64 numpy restatement of the warm-start network (TEST INFRASTRUCTURE ONLY).

The checker for the GPU network's forward pass and gradients (libfehip's
fehip_mlp_*).  Same layer shapes and flat parameter layout:
W0 [16x128], b0 [128], W1 [128x128], b1, W2 [128x128], b2, W3 [128x16], b3 [16];
ReLU on the first three layers; prediction = column 0 of the last layer;
loss = 0.5 * mean (y - t)^2.

``emulate_bf16=True`` rounds exactly the operands the GPU rounds to bf16
(inputs, weights, stored activations, the loss gradient and the back-propagated
gradients) and keeps everything else in FP64, so the remaining difference is
FP32 accumulation order only.  ``emulate_bf16=False`` is the exact FP64
network.  Pinned by central finite differences in tests/test_mlp_oracle.py.
"""
import numpy as np

D_IN = (16, 128, 128, 128)
D_OUT = (128, 128, 128, 16)


def bf16(x):
    """Round to bfloat16 (nearest even), returned as float64 values."""
    a = np.ascontiguousarray(np.asarray(x, np.float32)).view(np.uint32).astype(np.uint64)
    a = (a + 0x7FFF + ((a >> 16) & 1)) >> 16 << 16
    return a.astype(np.uint32).view(np.float32).astype(np.float64)


def layout():
    offs, o = [], 0
    for din, dout in zip(D_IN, D_OUT):
        offs.append((o, o + din * dout))
        o += din * dout + dout
    return offs, o


def unpack(params):
    offs, n = layout()
    p = np.asarray(params, np.float64)
    assert p.size == n
    Ws, bs = [], []
    for (ow, ob), din, dout in zip(offs, D_IN, D_OUT):
        Ws.append(p[ow:ob].reshape(din, dout))
        bs.append(p[ob:ob + dout])
    return Ws, bs


def pack(Ws, bs):
    return np.concatenate([np.concatenate([W.ravel(), b]) for W, b in zip(Ws, bs)])


def forward(params, X, emulate_bf16=False):
    Ws, bs = unpack(params)
    r = bf16 if emulate_bf16 else (lambda a: np.asarray(a, np.float64))
    H = [r(X)]
    Z = []
    for l in range(4):
        z = H[l] @ r(Ws[l]) + bs[l]
        Z.append(z)
        if l < 3:
            H.append(r(np.maximum(z, 0.0)))
    return Z[3][:, 0], H, Z


def loss_and_grad(params, X, t, emulate_bf16=False):
    Ws, bs = unpack(params)
    r = bf16 if emulate_bf16 else (lambda a: np.asarray(a, np.float64))
    y, H, Z = forward(params, X, emulate_bf16)
    n = len(t)
    d = y - np.asarray(t, np.float64)
    loss = 0.5 * np.mean(d * d)
    dZ = np.zeros((n, 16))
    dZ[:, 0] = d / n
    dZ = r(dZ)
    gW, gb = [None] * 4, [None] * 4
    for l in range(3, -1, -1):
        gW[l] = H[l].T @ dZ
        gb[l] = dZ.sum(axis=0)
        if l > 0:
            dH = dZ @ r(Ws[l]).T
            dZ = r(dH * (Z[l - 1] > 0))
    return loss, pack(gW, gb)
