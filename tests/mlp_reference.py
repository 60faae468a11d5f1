"""FP64 reference of the surrogate network (test infrastructure): forward and backward of a
GELU (erf form) fully connected network with mean-squared-error loss, optionally rounding weights,
activations and the backward signals to bf16 (round to nearest even) at the points where the GPU
network holds them in bf16: layer inputs, weights, the output gradient and the gradients at the
pre-activations.  Everything else is FP64."""
import numpy as np
from scipy.special import erf


def bf16(x):
    """Round to bf16 (nearest even) and return as float64."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def forward(params, x, emulate=False):
    q = bf16 if emulate else (lambda a: np.asarray(a, np.float64))
    h = q(x)
    hs, zs = [], []
    for l, (w, b) in enumerate(params):
        hs.append(h)
        z = h @ q(w).T + np.asarray(b, np.float64)
        zs.append(z)
        h = z if l == len(params) - 1 else q(gelu(z))
    return h, (hs, zs)


def loss_and_grads(params, x, t, emulate=False):
    q = bf16 if emulate else (lambda a: np.asarray(a, np.float64))
    y, (hs, zs) = forward(params, x, emulate)
    t = np.asarray(t, np.float64)
    n, m = y.shape
    loss = float(np.mean((y - t) ** 2))
    g = q(2.0 * (y - t) / (n * m))
    grads = [None] * len(params)
    for l in range(len(params) - 1, -1, -1):
        grads[l] = (g.T @ hs[l], g.sum(axis=0))
        if l == 0:
            break
        g = q((g @ q(params[l][0])) * gelu_grad(zs[l - 1]))
    return loss, grads
