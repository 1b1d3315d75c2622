"""Shared builders for parity tests: closed-form flows sampled in pyFTLE's layouts."""

import numpy as np

from oracle import flows
from oracle import numpy_oracle as O


def double_gyre_window(nx=64, ny=32, n_snap=11, snap_dt=0.1, t0=0.0, **kw):
    """Snapshots (list of (2, nx*ny)) and coordinates (2, nx*ny) of the double gyre."""
    x, y = flows.grid_axes_2d(nx, ny)
    vels = []
    pts = None
    for k in range(n_snap):
        v, p = flows.sample_2d(flows.double_gyre, t0 + k * snap_dt, x, y, **kw)
        vels.append(v)
        pts = p
    return vels, pts


def abc_window(n=16, n_snap=1, snap_dt=0.1, unsteady=False, endpoint=False):
    L = 2 * np.pi
    ax = np.linspace(0.0, L, n, endpoint=endpoint)
    vels = []
    pts = None
    for k in range(n_snap):
        v, p = flows.sample_3d(lambda t, q: flows.abc(t, q, unsteady=unsteady), k * snap_dt, ax, ax, ax)
        vels.append(v)
        pts = p
    return vels, pts


def seeds_2d(nsx=20, nsy=10, delta=1e-3, lo=(0.1, 0.1), hi=(1.9, 0.9)):
    xs = np.linspace(lo[0], hi[0], nsx)
    ys = np.linspace(lo[1], hi[1], nsy)
    X, Y = np.meshgrid(xs, ys)
    c = np.stack([X.ravel(), Y.ravel()], axis=1)
    return O.make_neighbor_positions(c, delta)


def seeds_3d(ns=6, delta=1e-3, lo=0.5, hi=5.5):
    s = np.linspace(lo, hi, ns)
    Z, Y, X = np.meshgrid(s, s, s, indexing="ij")
    c = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    return O.make_neighbor_positions(c, delta)


def grid_fields(grid_shape, vels, pts, periodic=None):
    return [O.GridField(grid_shape, v, pts, periodic) for v in vels]


# ---- torch restatements of the closed-form flows: the device-callable velocity_fn
# a user passes to AnalyticalSolver (the same formulas as oracle/flows.py) ----

def torch_double_gyre(t, points, A=0.1, eps=0.25, omega=2.0 * np.pi / 10.0):
    import torch

    x, y = points[:, 0], points[:, 1]
    s = np.sin(omega * t)
    f = eps * s * x**2 + (1.0 - 2.0 * eps * s) * x
    dfdx = 2.0 * eps * s * x + (1.0 - 2.0 * eps * s)
    u = -np.pi * A * torch.sin(np.pi * f) * torch.cos(np.pi * y)
    v = np.pi * A * torch.cos(np.pi * f) * torch.sin(np.pi * y) * dfdx
    return torch.stack([u, v], dim=1)


def torch_helical_vortex(t, points, Lz=8.0, R0=1.0, Gamma=5.0, sigma=0.3, omega=2.0 * np.pi / 2.0):
    import torch

    x, y, z = points[:, 0], points[:, 1], points[:, 2]
    k = 8.0 * np.pi / Lz
    phi = omega * t
    xr = x - R0 * torch.cos(k * z + phi)
    yr = y - R0 * torch.sin(k * z + phi)
    r = torch.sqrt(xr**2 + yr**2)
    theta = torch.atan2(yr, xr)
    v_theta = Gamma / (2.0 * np.pi * r) * (1.0 - torch.exp(-((r / sigma) ** 2)))
    u = -v_theta * torch.sin(theta)
    v = v_theta * torch.cos(theta)
    w = 0.2 * Gamma / (2.0 * np.pi * R0) * torch.exp(-((r / sigma) ** 2))
    scale = torch.max(torch.sqrt(u**2 + v**2 + w**2))
    return torch.stack([u / scale, v / scale, w / scale], dim=1)


def helical_particles(nxp=8, nyp=8, nzp=16, Lxp=3.8, Lyp=3.8, Lzp=7.6):
    """examples/01_Closed-form_velocity_helical.ipynb (particles cell): seed grid,
    delta = half the smallest half-spacing, blocks [L, R, T, B, F, Bk]."""
    xp = np.linspace(-Lxp / 2, Lxp / 2, nxp)
    yp = np.linspace(-Lyp / 2, Lyp / 2, nyp)
    zp = np.linspace(0, Lzp, nzp)
    Xc, Yc, Zc = np.meshgrid(xp, yp, zp, indexing="ij")
    dx, dy, dz = (np.min(np.gradient(a)) / 2 for a in (xp, yp, zp))
    delta = 0.5 * np.min((dx, dy, dz))
    c = np.column_stack((Xc.ravel(), Yc.ravel(), Zc.ravel()))
    return O.make_neighbor_positions(c, delta)
