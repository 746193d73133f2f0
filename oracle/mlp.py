"""CPU oracle of the lambda* warm-start network -- TEST INFRASTRUCTURE ONLY.

The network is a 7 -> 64 -> 64 -> 64 -> 2 multilayer perceptron with SiLU on
the hidden layers.  Its parameters are one flat float32 vector:

    W1 (64x7), b1 (64), W2 (64x64), b2 (64), W3 (64x64), b3 (64), W4 (2x64), b4 (2)

with W stored row-major as (out, in), so a layer is z = W a + b.

``forward``/``loss_grad`` evaluate it in FP64.  With ``bf16=True`` they round
exactly the operands that the device rounds before its BF16 matrix-core
products: the weights, each layer's input activations and the back-propagated
gradients that enter a product.  Biases, pre-activations, SiLU and all sums
stay wide, as FP32 accumulation does on the device.  The device and this
oracle therefore differ only by FP32-vs-FP64 accumulation and the rare
operand whose rounding lands on the other side of a BF16 boundary.
"""
from __future__ import annotations

import numpy as np

SIZES = [(64, 7), (64, 64), (64, 64), (2, 64)]
NPARAM = sum(o * i + o for o, i in SIZES)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    out, k = [], 0
    for o, i in SIZES:
        W = p[k:k + o * i].reshape(o, i); k += o * i
        b = p[k:k + o]; k += o
        out.append((W, b))
    return out


def pack(layers):
    return np.concatenate([np.concatenate([W.ravel(), b]) for W, b in layers]).astype(np.float32)


def bf16(x):
    """Round to bfloat16 (nearest, ties to even), returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(p, X, bf=True):
    """Outputs (n, 2) for inputs (n, 7)."""
    q = bf16 if bf else (lambda v: np.asarray(v, dtype=np.float64))
    a = q(X)
    L = unpack(p)
    for k, (W, b) in enumerate(L):
        z = a @ q(W).T + b
        if k == len(L) - 1:
            return z
        a = q(silu(z))


def loss_grad(p, X, T, w=(1.0, 1.0), bf=True):
    """Loss sum_o w_o (y_o - t_o)^2 averaged over the batch, and its gradient
    with respect to the flat parameter vector (float64)."""
    q = bf16 if bf else (lambda v: np.asarray(v, dtype=np.float64))
    L = unpack(p)
    n = X.shape[0]
    acts, zs = [q(X)], []
    a = acts[0]
    for k, (W, b) in enumerate(L):
        z = a @ q(W).T + b
        zs.append(z)
        if k < len(L) - 1:
            a = q(silu(z))
            acts.append(a)
    y = zs[-1]
    wv = np.asarray(w, dtype=np.float64)
    loss = float(np.sum(wv * (y - T) ** 2) / n)
    dz = 2.0 * wv * (y - T) / n
    grads = [None] * len(L)
    for k in range(len(L) - 1, -1, -1):
        W, _ = L[k]
        g = q(dz)
        grads[k] = (g.T @ acts[k], dz.sum(0))
        if k > 0:
            dh = g @ q(W)
            dz = dh * dsilu(zs[k - 1])
    return loss, np.concatenate([np.concatenate([dW.ravel(), db]) for dW, db in grads])


def adam(p, m, v, g, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba, Algorithm 1) on float64 copies; t >= 1."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v
