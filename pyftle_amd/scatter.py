"""Scattered-data fields (interpolate.py:188-294): the device-side structures behind
``LinearInterpolator`` (SciPy ``LinearNDInterpolator``) and
``NearestNeighborInterpolator`` (SciPy ``NearestNDInterpolator``).

``ScatterData(points)`` triangulates the data points with Qhull through
``scipy.spatial.Delaunay`` -- what ``LinearNDInterpolator`` itself does, with the
same default options -- and keeps its ``transform`` (per simplex the inverse
barycentric map and reference vertex) and hull flags.  Two uniform bucket grids
(about two simplices, or two points, per bucket) index the simplices by their
bounding boxes and the points by position, as CSR lists.  Everything is uploaded
once per point set; the velocities of each snapshot are one record per point in
the device ring, and the kernels locate, blend and search on the GPU.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _native as N
from .field import cuda_device


def _grid_for(lo, hi, n_items, per_bucket=2.0):
    """Uniform bucket grid over [lo, hi] holding about ``per_bucket`` items per bucket."""
    ext = np.maximum(hi - lo, 1e-12 * max(1.0, float(np.abs(np.concatenate([lo, hi])).max())))
    dim = len(lo)
    cell = (float(np.prod(ext)) * per_bucket / max(1, n_items)) ** (1.0 / dim)
    dims = np.clip(np.ceil(ext / cell), 1, 4096).astype(np.int64)
    while int(np.prod(dims)) >= 1 << 30:
        dims = np.maximum(1, dims // 2)
    return lo, dims, ext / dims


def _csr(bucket, ids, n_buckets):
    order = np.argsort(bucket, kind="stable")
    offsets = np.zeros(n_buckets + 1, dtype=np.int64)
    np.cumsum(np.bincount(bucket, minlength=n_buckets), out=offsets[1:])
    return offsets.astype(np.int32), ids[order].astype(np.int32)


def ct_factors(P, simplices, transform, neighbors):
    """Cross-edge direction factors of the Clough-Tocher patches, (S, 3): for the edge
    opposite vertex k, from the barycentric coordinates c of the neighbouring
    triangle's centroid, g = (2 c_{k+2} + c_{k+1} - 1) / (2 - 3 c_{k+2} - 3 c_{k+1})
    (indices mod 3), and -1/2 where the edge lies on the hull."""
    S = simplices.shape[0]
    g = np.full((S, 3), -0.5)
    for k in range(3):
        nb = neighbors[:, k]
        has = nb >= 0
        y = P[simplices[nb[has]]].mean(axis=1)  # neighbour centroids
        T = transform[has]
        cc = np.einsum("sij,sj->si", T[:, :2, :], y - T[:, 2, :])
        c = np.column_stack([cc, 1.0 - cc.sum(axis=1)])
        a, b = c[:, (k + 2) % 3], c[:, (k + 1) % 3]
        g[has, k] = (2.0 * a + b - 1.0) / (2.0 - 3.0 * a - 3.0 * b)
    return g


def build_host(points, need_simplices: bool = True) -> dict:
    """The host structures of a point set ``(dim, N)`` as NumPy arrays: points
    ``(N, dim)``; point bucket grid (``p_org``, ``p_dims``, ``p_size``) with CSR
    ``p_off``/``p_ids``; and, with ``need_simplices``, the Delaunay simplices,
    ``transform``, hull flags and the simplex bucket grid with CSR ``s_off``/``s_ids``."""
    pts = points.detach().cpu().numpy() if isinstance(points, torch.Tensor) else np.asarray(points)
    pts = np.ascontiguousarray(pts, dtype=np.float64)
    if pts.ndim != 2 or pts.shape[0] not in (2, 3):
        raise ValueError("points must have shape (dim, N) with dim 2 or 3")
    dim, n = pts.shape
    P = np.ascontiguousarray(pts.T)
    out = {"dim": dim, "n_points": n, "points": P}
    lo, hi = P.min(0), P.max(0)
    pad = 1e-7 * np.maximum(hi - lo, 1e-300) + 1e-300
    lo, hi = lo - pad, hi + pad
    org, dims, size = _grid_for(lo, hi, n)
    k = np.clip(np.floor((P - org) / size).astype(np.int64), 0, dims - 1)
    flat = k[:, 0] + dims[0] * (k[:, 1] + (dims[1] * k[:, 2] if dim == 3 else 0))
    out["p_off"], out["p_ids"] = _csr(flat, np.arange(n), int(np.prod(dims)))
    out["p_org"], out["p_dims"], out["p_size"] = org, dims, size
    if need_simplices:
        from scipy.spatial import Delaunay

        tri = Delaunay(P)  # LinearNDInterpolator's own call, default Qhull options
        simp = np.ascontiguousarray(tri.simplices, dtype=np.int32)
        S = simp.shape[0]
        out["tri"] = tri
        out["simplices"] = simp
        out["transform"] = np.ascontiguousarray(tri.transform, dtype=np.float64)
        out["hull"] = np.ascontiguousarray((tri.neighbors < 0).any(axis=1).astype(np.int32))
        org, dims, size = _grid_for(lo, hi, S)
        V = P[simp]  # (S, dim+1, dim)
        bmin, bmax = V.min(axis=1), V.max(axis=1)
        grow = 1e-7 * (bmax - bmin).max(axis=1, keepdims=True) + 1e-300
        i0 = np.clip(np.floor((bmin - grow - org) / size).astype(np.int64), 0, dims - 1)
        i1 = np.clip(np.floor((bmax + grow - org) / size).astype(np.int64), 0, dims - 1)
        span = i1 - i0 + 1
        counts = np.prod(span, axis=1)
        sid = np.repeat(np.arange(S), counts)
        local = np.arange(int(counts.sum())) - np.repeat(np.cumsum(counts) - counts, counts)
        bx = i0[sid, 0] + local % span[sid, 0]
        rest = local // span[sid, 0]
        by = i0[sid, 1] + rest % span[sid, 1]
        flat = bx + dims[0] * by
        if dim == 3:
            bz = i0[sid, 2] + rest // span[sid, 1]
            flat = flat + dims[0] * dims[1] * bz
        out["s_off"], out["s_ids"] = _csr(flat, sid, int(np.prod(dims)))
        out["s_org"], out["s_dims"], out["s_size"] = org, dims, size
    return out


class ScatterData:
    """Triangulation, hull flags and bucket grids of one point set, on the device
    (``cubic``: also the Clough-Tocher cross-edge factors; ``tri`` keeps SciPy's
    triangulation on the host for the gradient estimate)."""

    def __init__(self, points, need_simplices: bool = True, cubic: bool = False):
        h = build_host(points, need_simplices or cubic)
        self.dim, self.n_points = h["dim"], h["n_points"]
        self.tri = h.get("tri")
        self._keep = []
        d = N.Scatter()
        d.n_points = self.n_points
        d.points = self._upload(h["points"])
        d.point_offsets = self._upload(h["p_off"])
        d.point_ids = self._upload(h["p_ids"])
        for a in range(self.dim):
            d.point_dims[a] = int(h["p_dims"][a])
            d.point_origin[a] = float(h["p_org"][a])
            d.point_size[a] = float(h["p_size"][a])
        self.n_simplices = 0
        if cubic:
            tri = h["tri"]
            d.ct_g = self._upload(ct_factors(h["points"], tri.simplices, tri.transform, tri.neighbors))
        if need_simplices or cubic:
            self.n_simplices = d.n_simplices = h["simplices"].shape[0]
            d.simplices = self._upload(h["simplices"])
            d.transform = self._upload(h["transform"])
            d.hull = self._upload(h["hull"])
            d.simplex_offsets = self._upload(h["s_off"])
            d.simplex_ids = self._upload(h["s_ids"])
            for a in range(self.dim):
                d.simplex_dims[a] = int(h["s_dims"][a])
                d.simplex_origin[a] = float(h["s_org"][a])
                d.simplex_size[a] = float(h["s_size"][a])
        self.desc = d

    def _upload(self, arr):
        t = torch.from_numpy(np.ascontiguousarray(arr)).to(cuda_device())
        self._keep.append(t)
        return t.data_ptr()

    @property
    def address(self) -> int:
        return ctypes.addressof(self.desc)
