"""Float64 restatement of the surrogate MLP (test infrastructure only).

Forward: x' = (x - shift) * scale; h_l = GELU(h_{l-1} W_l^T + b_l) for the hidden layers
(GELU in its tanh form, 0.5 x (1 + tanh(sqrt(2/pi) (x + 0.044715 x^3))), the activation the
device network uses); y = h W_L^T + b_L.
Loss: L = sum W (y - t)^2 / sum W.  Gradients by reverse mode.  Parameters use the
library's flat layout: per layer W [out][in] row-major, then b [out].

`round_bf16` reproduces the device operand precision (round to nearest even to 8
significant bits) so a gradient check can compare like with like.  The backward pass
itself is pinned by central finite differences in the tests.
"""
from __future__ import annotations

import math

import numpy as np


def layer_shapes(n_in=4, hidden=256, n_hidden_layers=3, n_out=128):
    dims = [n_in] + [hidden] * n_hidden_layers + [n_out]
    return [(dims[i + 1], dims[i]) for i in range(len(dims) - 1)]


def unflatten(p, shapes):
    out, o = [], 0
    for (n_o, n_i) in shapes:
        W = p[o:o + n_o * n_i].reshape(n_o, n_i)
        o += n_o * n_i
        b = p[o:o + n_o]
        o += n_o
        out.append((W, b))
    assert o == p.size
    return out


def round_bf16(x):
    """float -> bfloat16 (round to nearest even) -> float64."""
    a = np.asarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


_C = math.sqrt(2.0 / math.pi)


def gelu(z):
    return 0.5 * z * (1.0 + np.tanh(_C * (z + 0.044715 * z ** 3)))


def gelu_grad(z):
    t = np.tanh(_C * (z + 0.044715 * z ** 3))
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * _C * (1.0 + 3 * 0.044715 * z * z)


def forward(p, X, shapes, shift=None, scale=None, bf16=False):
    """Returns (y, cache).  bf16=True rounds weights, inputs and stored activations the
    way the device kernels hold them."""
    params = unflatten(np.asarray(p, np.float64), shapes)
    x = np.asarray(X, np.float64)
    if shift is not None:
        x = (x - shift) * scale
    rnd = round_bf16 if bf16 else (lambda a: np.asarray(a, np.float64))
    h = rnd(x)
    cache = [h]
    for li, (W, b) in enumerate(params):
        z = h @ rnd(W).T + b
        if li + 1 < len(params):
            cache.append(z)
            h = rnd(gelu(z))
            cache.append(h)
        else:
            y = z
    return y, cache


def loss_and_grad(p, X, T, Wt, shapes, shift=None, scale=None, bf16=False, n_pois=0, R0=None):
    """Loss normalized by sum(W): Poisson NLL w (r - t log r) of the rate r = e^y (+ R0 when
    rate offsets are given) on the first n_pois outputs, w (y - t)^2 on the rest; and its
    gradient."""
    params = unflatten(np.asarray(p, np.float64), shapes)
    y, cache = forward(p, X, shapes, shift, scale, bf16)
    w = np.ones_like(y) if Wt is None else np.asarray(Wt, np.float64)
    ws = w.sum()
    t = np.asarray(T, np.float64)
    d = y - t
    lmat = w * d * d
    gmat = 2.0 * w * d
    if n_pois:
        e = np.exp(y[:, :n_pois])
        if R0 is None:
            lmat[:, :n_pois] = w[:, :n_pois] * (e - t[:, :n_pois] * y[:, :n_pois])
            gmat[:, :n_pois] = w[:, :n_pois] * (e - t[:, :n_pois])
        else:
            r = e + np.asarray(R0, np.float64)
            lmat[:, :n_pois] = w[:, :n_pois] * (r - t[:, :n_pois] * np.log(r))
            gmat[:, :n_pois] = w[:, :n_pois] * (1.0 - t[:, :n_pois] / r) * e
    L = float(lmat.sum() / ws)
    rnd = round_bf16 if bf16 else (lambda a: np.asarray(a, np.float64))
    dz = gmat / ws
    grads = []
    n = len(params)
    for li in range(n - 1, -1, -1):
        W, b = params[li]
        h_in = cache[2 * li]
        gW = dz.T @ h_in
        gb = dz.sum(0)
        grads.append((gW, gb))
        if li == 0:
            break
        z_prev = cache[2 * li - 1]
        dz = (dz @ rnd(W)) * gelu_grad(z_prev)
    grads.reverse()
    flat = np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])
    return L, flat
