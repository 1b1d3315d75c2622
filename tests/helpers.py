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
