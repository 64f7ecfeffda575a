"""Test infrastructure: the float64 NumPy oracle of the warm-start network (no reference
counterpart; the reference has no learned component).  Only tests/, smoke() and bench.py's
CPU baseline may use it, and only as the checker.

The network: n_layers 'same'-padded 1D convolutions over the L grid points of a sample,
Y[b, co, x] = Σ_{ci, t} W[co, ci, t] X[b, ci, x + t - h] + bias[co] (h = kernel // 2, taps outside
0..L-1 read as zero), SiLU z·σ(z) after every layer but the last; loss = mean((y - target)²).
Parameters are packed per layer as W [cout, cin, kernel] (C order) then bias [cout].

Pinned by tests/test_net_oracle.py: the convolution against numpy.correlate on a single channel,
and the gradient against central finite differences in float64.
"""
import numpy as np


def layout(n_layers, kernel, c_in, c_hidden, c_out):
    """[(cin, cout, w_offset, b_offset)] and the parameter count."""
    out, o = [], 0
    for l in range(n_layers):
        ci = c_in if l == 0 else c_hidden
        co = c_out if l == n_layers - 1 else c_hidden
        wo = o
        o += co * ci * kernel
        out.append((ci, co, wo, o))
        o += co
    return out, o


def _unpack(p, cfg):
    lay, _ = layout(*cfg)
    k = cfg[1]
    return [(p[wo:wo + co * ci * k].reshape(co, ci, k), p[bo:bo + co]) for ci, co, wo, bo in lay]


def _conv(X, W):
    """Y[b, o, x] = Σ_{c,t} W[o, c, t] Xpad[b, c, x + t]."""
    n, ci, L = X.shape
    k = W.shape[2]
    h = k // 2
    Xp = np.zeros((n, ci, L + 2 * h))
    Xp[:, :, h:h + L] = X
    Y = np.zeros((n, W.shape[0], L))
    for t in range(k):
        Y += np.einsum("oc,bcx->box", W[:, :, t], Xp[:, :, t:t + L])
    return Y, Xp


def forward(p, x, cfg):
    """y [n, c_out, L] for x [n, c_in, L]; cfg = (n_layers, kernel, c_in, c_hidden, c_out)."""
    X = np.asarray(x, dtype=np.float64)
    layers = _unpack(np.asarray(p, dtype=np.float64), cfg)
    for l, (W, b) in enumerate(layers):
        Z = _conv(X, W)[0] + b[None, :, None]
        X = Z * (1 / (1 + np.exp(-Z))) if l < len(layers) - 1 else Z
    return X


def loss_and_grad(p, x, target, cfg, _plant_silu_error=False):
    """mean((y - target)²) and its gradient with respect to the packed parameters.
    `_plant_silu_error` drops the z (1 - s) term of SiLU' (for the test that the gradient
    check detects such a mistake)."""
    p = np.asarray(p, dtype=np.float64)
    layers = _unpack(p, cfg)
    X = np.asarray(x, dtype=np.float64)
    cache = []
    for l, (W, b) in enumerate(layers):
        Y, Xp = _conv(X, W)
        Z = Y + b[None, :, None]
        cache.append((Xp, Z))
        X = Z * (1 / (1 + np.exp(-Z))) if l < len(layers) - 1 else Z
    d = X - np.asarray(target, dtype=np.float64)
    loss = np.mean(d * d)
    dA = 2 * d / d.size
    grad = np.zeros_like(p)
    lay, _ = layout(*cfg)
    k = cfg[1]
    h = k // 2
    for l in range(len(layers) - 1, -1, -1):
        W, b = layers[l]
        Xp, Z = cache[l]
        if l < len(layers) - 1:
            s = 1 / (1 + np.exp(-Z))
            dZ = dA * (s if _plant_silu_error else s * (1 + Z * (1 - s)))
        else:
            dZ = dA
        ci, co, wo, bo = lay[l]
        L = Z.shape[2]
        dW = np.zeros_like(W)
        dXp = np.zeros_like(Xp)
        for t in range(k):
            dW[:, :, t] = np.einsum("box,bcx->oc", dZ, Xp[:, :, t:t + L])
            dXp[:, :, t:t + L] += np.einsum("oc,box->bcx", W[:, :, t], dZ)
        grad[wo:wo + W.size] = dW.ravel()
        grad[bo:bo + co] = dZ.sum(axis=(0, 2))
        dA = dXp[:, :, h:h + L]
    return loss, grad
