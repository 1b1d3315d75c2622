"""Closed-form velocity fields used as oracle inputs (FP64, NumPy).

* double gyre: examples/01_Closed-form_velocity.ipynb (cell 1) and
  examples/02_File-based_IO.ipynb (double-gyre cell), ``velocity_fn(t, xy)``;
* ABC flow (steady, and the unsteady ``A(t)`` variant): textbook
  Arnold-Beltrami-Childress field, ``u = A sin z + C cos y``,
  ``v = B sin x + A cos z``, ``w = C sin y + B cos x``;
* saddle ``u = l x, v = -l y``: linear field, so bilinear interpolation and
  the flow map are exact (used for closed-form end-to-end checks);
* helical vortex filament: examples/01_Closed-form_velocity_helical.ipynb
  (cell 1), normalised by the largest speed over the queried points.

All functions take ``points`` of shape ``(M, dim)`` and return ``(M, dim)``.
Grid samplers return the reference's ``(dim, Ngrid)`` velocity and coordinate
arrays in the layout ``HighPerfInterpolator`` expects (interpolate.py:342-424).
"""

from __future__ import annotations

import numpy as np


def double_gyre(t, points, A=0.1, eps=0.25, omega=2.0 * np.pi / 10.0):
    x = points[:, 0]
    y = points[:, 1]
    s = np.sin(omega * t)
    f = eps * s * x**2 + (1.0 - 2.0 * eps * s) * x
    dfdx = 2.0 * eps * s * x + (1.0 - 2.0 * eps * s)
    out = np.empty_like(points, dtype=np.float64)
    out[:, 0] = -np.pi * A * np.sin(np.pi * f) * np.cos(np.pi * y)
    out[:, 1] = np.pi * A * np.cos(np.pi * f) * np.sin(np.pi * y) * dfdx
    return out


def abc(t, points, A=np.sqrt(3.0), B=np.sqrt(2.0), C=1.0, unsteady=False):
    x, y, z = points[:, 0], points[:, 1], points[:, 2]
    a = A + 0.5 * np.sin(np.pi * t / 5.0) if unsteady else A
    out = np.empty_like(points, dtype=np.float64)
    out[:, 0] = a * np.sin(z) + C * np.cos(y)
    out[:, 1] = B * np.sin(x) + a * np.cos(z)
    out[:, 2] = C * np.sin(y) + B * np.cos(x)
    return out


def saddle(t, points, lam=0.5):
    out = np.empty_like(points, dtype=np.float64)
    out[:, 0] = lam * points[:, 0]
    out[:, 1] = -lam * points[:, 1]
    if points.shape[1] == 3:
        out[:, 2] = 0.0 * points[:, 2]
    return out


def helical_vortex(t, points, Lz=8.0, R0=1.0, Gamma=5.0, sigma=0.3, omega=2.0 * np.pi / 2.0):
    """examples/01_Closed-form_velocity_helical.ipynb cell 1: swirl around a helical
    centreline rotating with phase omega*t, a weak axial part, divided by the
    largest speed among ``points`` (a reduction over the whole query)."""
    x, y, z = points[:, 0], points[:, 1], points[:, 2]
    k = 8.0 * np.pi / Lz
    phi = omega * t
    xr = x - R0 * np.cos(k * z + phi)
    yr = y - R0 * np.sin(k * z + phi)
    r = np.sqrt(xr**2 + yr**2)
    theta = np.arctan2(yr, xr)
    v_theta = Gamma / (2.0 * np.pi * r) * (1.0 - np.exp(-((r / sigma) ** 2)))
    u = -v_theta * np.sin(theta)
    v = v_theta * np.cos(theta)
    w = 0.2 * Gamma / (2.0 * np.pi * R0) * np.exp(-((r / sigma) ** 2))
    scale = np.max(np.sqrt(u**2 + v**2 + w**2))
    out = np.empty_like(points, dtype=np.float64)
    out[:, 0] = u / scale
    out[:, 1] = v / scale
    out[:, 2] = w / scale
    return out


def grid_axes_2d(nx, ny, xlim=(0.0, 2.0), ylim=(0.0, 1.0)):
    return np.linspace(xlim[0], xlim[1], nx), np.linspace(ylim[0], ylim[1], ny)


def sample_2d(fn, t, x, y, **kw):
    """Sample ``fn`` on ``meshgrid(x, y)`` flattened x-fastest.

    Returns ``(velocities (2, nx*ny), points (2, nx*ny))`` in the order of
    examples/02_File-based_IO.ipynb (``meshgrid`` + ``ravel``), which is the
    F-order layout interpolate.py:363 reshapes.
    """
    X, Y = np.meshgrid(x, y)
    pts = np.stack([X.ravel(), Y.ravel()], axis=1)
    vel = fn(t, pts, **kw)
    return np.ascontiguousarray(vel.T), np.ascontiguousarray(pts.T)


def sample_3d(fn, t, x, y, z, **kw):
    """Sample ``fn`` on ``meshgrid(x, y, z)`` ('xy' indexing) flattened C-order.

    This is the layout interpolate.py:383-385 untangles with its axis swap
    (consistent for ``nx == ny``, as in the reference).
    """
    X, Y, Z = np.meshgrid(x, y, z)
    pts = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    vel = fn(t, pts, **kw)
    return np.ascontiguousarray(vel.T), np.ascontiguousarray(pts.T)


def cell_centred_seeds_2d(nsx, nsy, xlim=(0.0, 2.0), ylim=(0.0, 1.0)):
    """Seeds at ``(i + 1/2) * Lx/nsx`` (row-major over y then x -> x fastest)."""
    xs = xlim[0] + (np.arange(nsx) + 0.5) * (xlim[1] - xlim[0]) / nsx
    ys = ylim[0] + (np.arange(nsy) + 0.5) * (ylim[1] - ylim[0]) / nsy
    X, Y = np.meshgrid(xs, ys)
    return np.stack([X.ravel(), Y.ravel()], axis=1)


def cell_centred_seeds_3d(ns, L=2.0 * np.pi):
    s = (np.arange(ns) + 0.5) * L / ns
    Z, Y, X = np.meshgrid(s, s, s, indexing="ij")
    return np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
