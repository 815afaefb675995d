"""FP64 NumPy restatement of the temperature-guess network (forward, loss,
backward) used to check the device training kernels' gradients.

TEST INFRASTRUCTURE ONLY (imported by tests/).  Parameter layout: W1[64][8],
b1[64], W2[64][64], b2[64], W3[64][64], b3[64], w4[64], b4 -- 8961 values.
Loss: mean over the batch of (y - t)^2.  Activation: silu(x) = x sigmoid(x).
"""
import numpy as np

NH, NIN = 64, 8


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    out = []
    for shape in ((NH, NIN), (NH,), (NH, NH), (NH,), (NH, NH), (NH,), (NH,), (1,)):
        n = int(np.prod(shape))
        out.append(p[o:o + n].reshape(shape))
        o += n
    return out


def silu(x):
    return x / (1.0 + np.exp(-x))


def dsilu(x):
    s = 1.0 / (1.0 + np.exp(-x))
    return s * (1.0 + x * (1.0 - s))


def forward(p, Z):
    W1, b1, W2, b2, W3, b3, w4, b4 = unpack(p)
    Z = np.asarray(Z, dtype=np.float64)
    a1 = Z @ W1.T + b1
    h1 = silu(a1)
    a2 = h1 @ W2.T + b2
    h2 = silu(a2)
    a3 = h2 @ W3.T + b3
    h3 = silu(a3)
    y = h3 @ w4 + b4[0]
    return y, (Z, a1, h1, a2, h2, a3, h3)


def loss_and_grad(p, Z, t):
    W1, b1, W2, b2, W3, b3, w4, b4 = unpack(p)
    y, (Z, a1, h1, a2, h2, a3, h3) = forward(p, Z)
    n = Z.shape[0]
    e = y - np.asarray(t, dtype=np.float64)
    loss = np.mean(e * e)
    dy = 2.0 * e / n
    dw4 = h3.T @ dy
    db4 = dy.sum()
    g3 = (dy[:, None] * w4[None, :]) * dsilu(a3)
    dW3 = g3.T @ h2
    db3 = g3.sum(axis=0)
    g2 = (g3 @ W3) * dsilu(a2)
    dW2 = g2.T @ h1
    db2 = g2.sum(axis=0)
    g1 = (g2 @ W2) * dsilu(a1)
    dW1 = g1.T @ Z
    db1 = g1.sum(axis=0)
    grad = np.concatenate([dW1.ravel(), db1, dW2.ravel(), db2, dW3.ravel(), db3, dw4, [db4]])
    return loss, grad


def bf16(x):
    """Round float32 values to bfloat16 (round to nearest even), returned as float64."""
    x = np.asarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def forward_bf16(p, Z):
    """The inference the device runs: BF16 weights and BF16 activations between
    layers, FP32-level accumulation (emulated in FP64), FP32 output layer."""
    W1, b1, W2, b2, W3, b3, w4, b4 = unpack(p)
    h = bf16(np.asarray(Z, dtype=np.float32)[:, :7])
    a = h @ bf16(W1[:, :7].astype(np.float32)).T + b1
    h = bf16(silu(a).astype(np.float32))
    a = h @ bf16(W2.astype(np.float32)).T + b2
    h = bf16(silu(a).astype(np.float32))
    a = h @ bf16(W3.astype(np.float32)).T + b3
    h = bf16(silu(a).astype(np.float32))
    return h @ w4 + b4[0]
