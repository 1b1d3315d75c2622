"""FP64 NumPy restatement of the flow-map surrogate MLP (test infrastructure).

Fourier features ``z0 = [sin(2 pi B v), cos(2 pi B v)]`` of ``v = x * scale``,
hidden layers ``h_{l+1} = silu(h_l W_l^T + b_l)``, output ``y = h W_out^T +
b_out``; backward pass of the MSE loss ``mean((y - t)^2)`` over all outputs;
forward-mode tangents (JVP) with respect to the inputs.  Pinned in
``tests/test_oracle_mlp.py`` by central finite differences in FP64 (gradient
and JVP) and by the closed form of a one-layer network.
"""

from __future__ import annotations

import numpy as np


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def fourier(x, Bm, scale):
    v = x * scale
    ph = 2.0 * np.pi * v @ Bm.T
    return np.concatenate([np.sin(ph), np.cos(ph)], axis=1)


def fourier_tangent(x, Bm, scale, a):
    v = x * scale
    ph = 2.0 * np.pi * v @ Bm.T
    dph = 2.0 * np.pi * Bm[:, a] * scale[a]
    return np.concatenate([np.cos(ph) * dph, -np.sin(ph) * dph], axis=1)


def forward(params, x, cache=False):
    """params: dict with B, scale, W (list), b (list), Wout, bout."""
    h = fourier(x, params["B"], params["scale"])
    hs, zs = [h], []
    for W, b in zip(params["W"], params["b"]):
        z = h @ W.T + b
        h = silu(z)
        zs.append(z)
        hs.append(h)
    y = h @ params["Wout"].T + params["bout"]
    if cache:
        return y, hs, zs
    return y


def loss_and_grads(params, x, t):
    """MSE = mean over all outputs; returns (loss, grads dict)."""
    y, hs, zs = forward(params, x, cache=True)
    n = y.size
    d = y - t
    loss = float(np.sum(d * d) / n)
    dy = 2.0 * d / n
    g = {"W": [None] * len(params["W"]), "b": [None] * len(params["b"])}
    g["Wout"] = dy.T @ hs[-1]
    g["bout"] = dy.sum(axis=0)
    dh = dy @ params["Wout"]
    for l in range(len(params["W"]) - 1, -1, -1):
        dz = dh * dsilu(zs[l])
        g["W"][l] = dz.T @ hs[l]
        g["b"][l] = dz.sum(axis=0)
        dh = dz @ params["W"][l]
    return loss, g


def jvp(params, x, a):
    """d y / d x_a (forward-mode), shape like y."""
    h = fourier(x, params["B"], params["scale"])
    th = fourier_tangent(x, params["B"], params["scale"], a)
    for W, b in zip(params["W"], params["b"]):
        z = h @ W.T + b
        tz = th @ W.T
        h = silu(z)
        th = tz * dsilu(z)
    return th @ params["Wout"].T


def init_params(rng, d_in=3, d_out=2, n_freq=8, width=16, depth=2, sigma=1.0, scale=None):
    sc = np.ones(d_in) if scale is None else np.asarray(scale, float)
    p = {"B": rng.standard_normal((n_freq, d_in)) * sigma, "scale": sc, "W": [], "b": []}
    fan = 2 * n_freq
    for _ in range(depth):
        p["W"].append(rng.standard_normal((width, fan)) * np.sqrt(2.0 / fan))
        p["b"].append(rng.standard_normal(width) * 0.1)
        fan = width
    p["Wout"] = rng.standard_normal((d_out, width)) * np.sqrt(1.0 / width)
    p["bout"] = rng.standard_normal(d_out) * 0.1
    return p
