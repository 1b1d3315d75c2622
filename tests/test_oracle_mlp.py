"""Pins of the MLP oracle (CPU, FP64): finite-difference gradients and JVP, closed form."""

import numpy as np

from oracle import mlp_oracle as M


def _fd_check(f, x, g, idx, h=1e-6):
    for i in idx:
        old = x.flat[i]
        x.flat[i] = old + h
        fp = f()
        x.flat[i] = old - h
        fm = f()
        x.flat[i] = old
        np.testing.assert_allclose(g.flat[i], (fp - fm) / (2 * h), rtol=2e-5, atol=1e-8)


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(3)
    p = M.init_params(rng, d_in=3, d_out=2, n_freq=5, width=7, depth=3, scale=(0.5, 1.0, 0.1))
    x = rng.uniform(0, 1, (11, 3))
    t = rng.standard_normal((11, 2))
    _, g = M.loss_and_grads(p, x, t)
    f = lambda: M.loss_and_grads(p, x, t)[0]  # noqa: E731
    for l in range(3):
        _fd_check(f, p["W"][l], g["W"][l], rng.choice(p["W"][l].size, 6, replace=False))
        _fd_check(f, p["b"][l], g["b"][l], range(p["b"][l].size))
    _fd_check(f, p["Wout"], g["Wout"], range(p["Wout"].size))
    _fd_check(f, p["bout"], g["bout"], range(2))


def test_jvp_matches_finite_differences():
    rng = np.random.default_rng(4)
    p = M.init_params(rng, d_in=3, d_out=2, n_freq=6, width=9, depth=2, sigma=2.0, scale=(0.5, 1.0, 0.1))
    x = rng.uniform(0, 1, (8, 3))
    h = 1e-6
    for a in range(3):
        e = np.zeros(3)
        e[a] = h
        fd = (M.forward(p, x + e) - M.forward(p, x - e)) / (2 * h)
        np.testing.assert_allclose(M.jvp(p, x, a), fd, rtol=1e-6, atol=1e-9)


def test_one_layer_closed_form():
    """With no hidden layer output = fourier(x) Wout^T + bout exactly."""
    rng = np.random.default_rng(5)
    p = M.init_params(rng, n_freq=4, width=3, depth=0)
    p["Wout"] = rng.standard_normal((2, 8))
    x = rng.uniform(0, 1, (5, 3))
    v = x * p["scale"]
    ph = 2 * np.pi * v @ p["B"].T
    exp = np.concatenate([np.sin(ph), np.cos(ph)], 1) @ p["Wout"].T + p["bout"]
    np.testing.assert_allclose(M.forward(p, x), exp, rtol=1e-14)
    assert abs(M.dsilu(0.0) - 0.5) < 1e-15 and abs(M.silu(0.0)) < 1e-15
