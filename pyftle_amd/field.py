"""Device-resident velocity fields: a ring of snapshot slots in HBM.

A ``DeviceField`` owns one HBM buffer holding ``ring_slots`` snapshot slots of
interleaved records (2D ``(u, v)``, 3D ``(u, v, w, 0)``; ``PAIR`` layout packs
snapshots ``k`` and ``k+1`` into one record so a time-linear lookup fetches
both levels in one load).  Node ``(ix, iy, iz)`` is record
``ix + nx*(iy + ny*iz)``.  Snapshots are copied in by the HIP pack kernel
from the component-planar arrays pyFTLE hands to ``Interpolator.update``
(interpolate.py:70-94, 446-472), whose flat layouts are expressed as strides:

* 2D: ``velocities.reshape((2, nx, ny), order="F")`` (interpolate.py:363),
  element ``(c, ix, iy)`` at ``c*N + ix + nx*iy``;
* 3D: ``transpose(velocities.reshape((3, nx, ny, nz)), (0, 2, 1, 3))``
  (interpolate.py:383-385); the C++ kernel then sees an ``(ny, nx, nz)``
  array, element ``(c, i, j, k)`` at ``c*N + j*ny*nz + i*nz + k``.

Grid geometry follows interpolate.py:359-401 and grid_interp.py:30-37:
axes ``linspace(min(coord), max(coord), n)`` and spacing
``(max - min)/(n - 1)``.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from . import _native as N

ALIGN = 256


def cuda_device() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("pyftle_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch.device("cuda", torch.cuda.current_device())


def stream_handle(stream: torch.cuda.Stream | None = None) -> ctypes.c_void_p:
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def dtype_code(dtype: torch.dtype) -> int:
    if dtype == torch.float32:
        return N.F32
    if dtype == torch.float64:
        return N.F64
    raise ValueError(f"unsupported dtype {dtype}; use float32 or float64")


def precision_dtype(precision) -> torch.dtype:
    if precision in (torch.float32, "float32", "f32", "fp32", np.float32):
        return torch.float32
    if precision in (torch.float64, "float64", "f64", "fp64", np.float64):
        return torch.float64
    raise ValueError(f"unknown precision {precision!r}")


def ptr(t: torch.Tensor) -> ctypes.c_void_p:
    return ctypes.c_void_p(t.data_ptr())


PINNED_MIN_BYTES = 1 << 20  # host arrays at least this large move through pinned buffers
_PINNABLE = {np.dtype(np.float32): torch.float32, np.dtype(np.float64): torch.float64,
             np.dtype(np.int32): torch.int32, np.dtype(np.int64): torch.int64}


_COPY_THREADS = 4
_copy_pool = None


def host_copy(dst: np.ndarray, src: np.ndarray) -> None:
    """``np.copyto(dst, src)``; large contiguous copies are split over a few
    threads (numpy releases the GIL for plain copies, and one core's memcpy
    falls well short of the host memory bandwidth)."""
    global _copy_pool
    if src.nbytes < (2 << 20) or dst.shape != src.shape or not (dst.flags.c_contiguous and
                                                                 src.flags.c_contiguous):
        np.copyto(dst, src, casting="same_kind")
        return
    if _copy_pool is None:
        from concurrent.futures import ThreadPoolExecutor

        _copy_pool = ThreadPoolExecutor(max_workers=_COPY_THREADS, thread_name_prefix="pyftle-copy")
    d = dst.reshape(-1)
    s_ = src.reshape(-1)
    cuts = np.linspace(0, d.size, _COPY_THREADS + 1).astype(np.int64)
    futs = [_copy_pool.submit(np.copyto, d[a:b], s_[a:b], "same_kind") for a, b in zip(cuts[:-1], cuts[1:])]
    for f in futs:
        f.result()


def to_device(x, dtype: torch.dtype | None = None) -> torch.Tensor:
    """Device copy of a tensor or host array.  Large host arrays are copied into a
    pinned buffer (torch's caching host allocator, which keeps the block until the
    transfer has run) and sent asynchronously; pageable transfers go through small
    driver staging chunks at a fraction of the bus rate."""
    dev = cuda_device()
    if isinstance(x, torch.Tensor):
        t = x.to(device=dev, dtype=dtype if dtype is not None else x.dtype)
    else:
        arr = np.asarray(x)
        tdt = _PINNABLE.get(arr.dtype)
        if tdt is not None and arr.nbytes >= PINNED_MIN_BYTES:
            host = torch.empty(arr.shape, dtype=tdt, pin_memory=True)
            host_copy(host.numpy(), arr)
            t = host.to(device=dev, non_blocking=True)
        else:
            t = torch.from_numpy(np.ascontiguousarray(arr)).to(device=dev)
        if dtype is not None:
            t = t.to(dtype)
    return t.contiguous()


def to_host(t: torch.Tensor, out: np.ndarray | None = None) -> np.ndarray:
    """Host copy of a device tensor; large ones come back through a pinned buffer
    (one DMA at the bus rate instead of the driver's pageable staging chunks).
    With ``out`` the values are written into that array (same shape) in place."""
    if t.numel() * t.element_size() >= PINNED_MIN_BYTES:
        host = torch.empty(t.shape, dtype=t.dtype, pin_memory=True)
        host.copy_(t, non_blocking=True)
        torch.cuda.current_stream().synchronize()
        h = host.numpy()
    else:
        h = t.cpu().numpy()
    if out is None:
        return h
    host_copy(out, h.reshape(out.shape))
    return out


@dataclass
class GridGeometry:
    """Array extents, origin and spacing per axis (device index order x, y, z)."""

    dim: int
    n: tuple
    origin: tuple
    spacing: tuple
    periodic: tuple
    src_strides: tuple  # element strides (c, x, y, z) of the reference flat layout
    grid_shape: tuple   # the user's grid_shape
    upper: tuple = ()   # last axis node (the RegularGridInterpolator modes)

    @staticmethod
    def from_reference(grid_shape, points, periodic=None) -> "GridGeometry":
        """Geometry of ``HighPerfInterpolator._initialize_interpolator``.

        ``points`` is the ``(dim, Ngrid)`` coordinate array of one snapshot
        (interpolate.py:359, 365-401).
        """
        grid_shape = tuple(int(s) for s in grid_shape)
        dim = len(grid_shape)
        if dim not in (2, 3):
            raise ValueError(f"Unsupported grid dimensionality: {dim}")
        pts = points.detach().cpu().numpy() if isinstance(points, torch.Tensor) else np.asarray(points)
        pts = pts.astype(np.float64, copy=False)
        ngrid = int(np.prod(grid_shape))
        if pts.shape != (dim, ngrid):
            raise ValueError(f"points must have shape {(dim, ngrid)}, got {pts.shape}")
        coords = pts.reshape((dim, *grid_shape))
        origin, spacing = [], []
        for a in range(dim):
            axis = np.linspace(np.min(coords[a]), np.max(coords[a]), grid_shape[a])
            origin.append(float(axis[0]))
            spacing.append(float((axis[-1] - axis[0]) / (axis.shape[0] - 1)))
        if dim == 2:
            nx, ny = grid_shape
            n = (nx, ny)
            strides = (ngrid, 1, nx, 0)
        else:
            nx, ny, nz = grid_shape
            n = (ny, nx, nz)  # extents of the transposed array the C++ kernel indexes
            strides = (ngrid, nz, ny * nz, 1)
        per = tuple(bool(p) for p in periodic) if periodic is not None else (False,) * dim
        return GridGeometry(dim, n, tuple(origin), tuple(spacing), per, strides, grid_shape)

    @staticmethod
    def from_rgi(grid_shape, points) -> "GridGeometry":
        """Geometry of ``GridInterpolator._initialize_interpolator`` (interpolate.py:598-660):
        axis ``a`` is ``numpy.linspace(points[a].min(), points[a].max(), n_a)`` and each
        velocity component is reshaped to ``grid_shape`` in C order (x slowest)."""
        grid_shape = tuple(int(s) for s in grid_shape)
        pts = points.detach().cpu().numpy() if isinstance(points, torch.Tensor) else np.asarray(points)
        pts = pts.astype(np.float64, copy=False)
        if pts.ndim != 2:
            raise ValueError("points must have shape (dim, Ngrid)")
        dim, n_points = pts.shape
        expected = int(np.prod(grid_shape))
        if expected != n_points:
            raise ValueError(f"grid_shape {grid_shape} implies {expected} points, but got {n_points}.")
        if dim not in (2, 3) or len(grid_shape) != dim:
            raise ValueError("Velocity field must have 2 or 3 components (u, v, [w]).")
        origin, spacing, upper = [], [], []
        for a in range(dim):
            axis = np.linspace(pts[a].min(), pts[a].max(), grid_shape[a])
            origin.append(float(axis[0]))
            upper.append(float(axis[-1]))
            spacing.append(float((axis[-1] - axis[0]) / (grid_shape[a] - 1)))
        if dim == 2:
            strides = (n_points, grid_shape[1], 1, 0)
        else:
            strides = (n_points, grid_shape[1] * grid_shape[2], grid_shape[2], 1)
        return GridGeometry(dim, grid_shape, tuple(origin), tuple(spacing), (False,) * dim, strides,
                            grid_shape, tuple(upper))

    @staticmethod
    def scattered(dim: int, n_points: int) -> "GridGeometry":
        """Record layout of a scattered-data field: one record per data point, the
        snapshot's ``(dim, N)`` components read point by point."""
        n = (int(n_points),) + (1,) * (dim - 1)
        return GridGeometry(dim, n, (0.0,) * dim, (1.0,) * dim, (False,) * dim, (int(n_points), 1, 0, 0),
                            (int(n_points),))

    @staticmethod
    def uniform(n, origin, spacing, periodic=None) -> "GridGeometry":
        """Geometry of a grid given directly in device order (x fastest).

        Snapshots for it are ``(dim, nz, ny, nx)`` component-planar arrays
        (or ``(dim, ny, nx)`` in 2D), i.e. flat ``ix + nx*(iy + ny*iz)``.
        """
        n = tuple(int(v) for v in n)
        dim = len(n)
        ngrid = int(np.prod(n))
        if dim == 2:
            strides = (ngrid, 1, n[0], 0)
        else:
            strides = (ngrid, 1, n[0], n[0] * n[1])
        per = tuple(bool(p) for p in periodic) if periodic is not None else (False,) * dim
        return GridGeometry(dim, n, tuple(float(o) for o in origin), tuple(float(s) for s in spacing),
                            per, strides, n)


class HostStager:
    """Pinned, double-buffered host -> device staging of snapshots.

    ``stage(array)`` copies a host snapshot into pinned buffer ``b`` (a host
    memcpy) and enqueues ``b``'s device copy on the caller's stream, so the file
    reading and memcpy of snapshot k+1 overlap the transfer and packing of
    snapshot k.  ``release(b)`` records, on the caller's stream, the event that
    marks buffer ``b`` free again once the work that reads it has run.  (A
    separate copy stream, joined by an event per snapshot, measured slower: the
    transfer is short next to the host memcpy, and each cross-stream wait costs
    tens of microseconds.)
    """

    def __init__(self, n_elems: int, dtype: torch.dtype, device: torch.device, buffers: int = 2):
        self.host = [torch.empty(n_elems, dtype=dtype, pin_memory=True) for _ in range(buffers)]
        self.dev = [torch.empty(n_elems, dtype=dtype, device=device) for _ in range(buffers)]
        self.free = [None] * buffers
        self.next = 0

    def stage(self, array: np.ndarray, stream: torch.cuda.Stream | None = None):
        b = self.next
        self.next = (b + 1) % len(self.host)
        if self.free[b] is not None:
            self.free[b].synchronize()  # the previous copy into dev[b] and its reader have run
        n = array.size
        host_copy(self.host[b][:n].numpy().reshape(array.shape), array)
        cur = stream if stream is not None else torch.cuda.current_stream()
        with torch.cuda.stream(cur):
            self.dev[b][:n].copy_(self.host[b][:n], non_blocking=True)
        return b, self.dev[b][:n]

    def release(self, b: int, stream: torch.cuda.Stream | None = None) -> None:
        ev = torch.cuda.Event()
        ev.record(stream if stream is not None else torch.cuda.current_stream())
        self.free[b] = ev


class DeviceField:
    """Ring of snapshot slots resident in HBM plus its C descriptor."""

    def __init__(self, geom: GridGeometry, *, dtype=torch.float32, layout=N.LAYOUT_SINGLE,
                 ring_slots: int = 1, n_snapshots: int = 1, t0: float = 0.0, dt_snap: float = 1.0,
                 interp: int = N.INTERP_GINTERP, scatter=None):
        self.geom = geom
        self.interp = int(interp)
        self.scatter = scatter  # ScatterData of the scattered-data modes (kept alive with the field)
        self.dtype = precision_dtype(dtype)
        self.layout = layout
        self.ring_slots = int(ring_slots)
        self.n_snapshots = int(n_snapshots)
        self.t0 = float(t0)
        self.dt_snap = float(dt_snap)
        dim = geom.dim
        levels = 2 if layout in (N.LAYOUT_PAIR, N.LAYOUT_CELL_PAIR) else 1
        cell = layout in (N.LAYOUT_CELL, N.LAYOUT_CELL_PAIR)
        # per-cell records hold the 2^dim corners of a cell (dim components each)
        comps = ((1 << dim) * dim if cell else (2 if dim == 2 else 4)) * levels
        if self.interp == N.INTERP_SCATTER_CUBIC:
            comps = 8  # value and gradient of u and v, padded
        esize = 4 if self.dtype == torch.float32 else 8
        # a periodic axis stores one ghost node (a copy of node 0) at index n; the cubic
        # mode stores spline coefficients -1..n per axis; the cell layouts n cells per axis
        if self.interp == N.INTERP_RGI_CUBIC:
            nodes = int(np.prod([geom.n[a] + 2 for a in range(dim)]))
        elif cell:
            nodes = int(np.prod([geom.n[a] for a in range(dim)]))
        else:
            nodes = int(np.prod([geom.n[a] + (1 if geom.periodic[a] else 0) for a in range(dim)]))
        self.record_bytes = comps * esize
        self.slot_stride = -(-nodes * self.record_bytes // ALIGN) * ALIGN
        total = self.slot_stride * self.ring_slots + ALIGN
        self._buf = torch.empty(total, dtype=torch.uint8, device=cuda_device())
        base = self._buf.data_ptr()
        self._offset = (-base) % ALIGN
        self.data_ptr = base + self._offset
        self.resident = {}  # slot -> snapshot index currently stored
        self._desc = self._make_desc()

    @property
    def nbytes(self) -> int:
        return self.slot_stride * self.ring_slots

    def _make_desc(self) -> N.Field:
        g = self.geom
        d = N.Field()
        d.dim = g.dim
        for a in range(3):
            d.n[a] = g.n[a] if a < g.dim else 1
            d.origin[a] = g.origin[a] if a < g.dim else 0.0
            d.spacing[a] = g.spacing[a] if a < g.dim else 1.0
            d.periodic[a] = int(g.periodic[a]) if a < g.dim else 0
        d.dtype = dtype_code(self.dtype)
        d.layout = self.layout
        d.ring_slots = self.ring_slots
        d.n_snapshots = self.n_snapshots
        d.slot_stride = self.slot_stride
        d.t0 = self.t0
        d.dt_snap = self.dt_snap
        d.data = self.data_ptr
        d.interp = self.interp
        for a in range(3):
            d.upper[a] = g.upper[a] if a < g.dim and len(g.upper) > a else 0.0
        d.scatter = self.scatter.address if self.scatter is not None else None
        return d

    @property
    def desc(self) -> N.Field:
        return self._desc

    def set_time_axis(self, n_snapshots: int, t0: float = 0.0, dt_snap: float = 1.0) -> None:
        self.n_snapshots = int(n_snapshots)
        self.t0 = float(t0)
        self.dt_snap = float(dt_snap)
        self._desc = self._make_desc()

    def slot_of(self, k: int) -> int:
        return k % self.ring_slots

    def write_records(self, k: int, records: np.ndarray) -> None:
        """Copy ready-made records ``(nodes, comps)`` of the field dtype into slot ``k``
        (the scattered cubic mode's values and gradients)."""
        rec = np.ascontiguousarray(records, dtype=np.float32 if self.dtype == torch.float32 else np.float64)
        nbytes = rec.nbytes
        if nbytes > self.slot_stride:
            raise ValueError(f"{nbytes} bytes of records exceed the slot ({self.slot_stride} bytes)")
        off = self._offset + self.slot_of(k) * self.slot_stride
        dst = self._buf.narrow(0, off, nbytes)
        src = to_device(rec.reshape(-1).view(np.uint8))
        dst.copy_(src, non_blocking=True)
        self._keepalive = src
        self.resident[self.slot_of(k)] = k

    def pack(self, k: int, velocities, strides=None, stream=None) -> None:
        """Copy snapshot ``k`` (component-planar, reference layout) into its slot.

        Host arrays of float32/float64 go through pinned double-buffered staging
        (``HostStager``); device tensors are packed directly.
        """
        if isinstance(velocities, np.ndarray) and velocities.dtype in (np.float32, np.float64) \
                and strides is None:
            tdt = torch.float64 if velocities.dtype == np.float64 else torch.float32
            st = getattr(self, "_stager", None)
            if st is None or st.host[0].numel() < velocities.size or st.host[0].dtype != tdt:
                st = self._stager = HostStager(velocities.size, tdt, cuda_device())
            b, src = st.stage(velocities, stream)
            self._pack_device(k, src, None, stream)
            st.release(b, stream)
            return
        src = velocities if isinstance(velocities, torch.Tensor) else to_device(velocities)
        if not src.is_cuda:
            src = src.to(cuda_device())
        self._pack_device(k, src, strides, stream)
        self._keepalive = src  # the copy is asynchronous: keep the source alive

    def _pack_device(self, k: int, src: torch.Tensor, strides, stream) -> None:
        if src.dtype not in (torch.float32, torch.float64):
            src = src.to(torch.float64)
        src = src.contiguous()
        st = strides if strides is not None else self.geom.src_strides
        n = list(self.geom.n) + [1] * (3 - self.geom.dim)
        extent = (self.geom.dim - 1) * st[0] + sum((n[a] - 1) * st[a + 1] for a in range(3)) + 1
        if any(s < 0 for s in st) or src.numel() < extent:
            raise ValueError(f"snapshot source has {src.numel()} elements; its strides need {extent}")
        arr = (ctypes.c_int64 * 4)(*[int(s) for s in st])
        if self.interp == N.INTERP_RGI_CUBIC:
            work = getattr(self, "_spline_work", None)
            if work is None:
                nbytes = int(N.load().pyftle_spline_work_bytes(ctypes.byref(self._desc)))
                work = self._spline_work = torch.empty(max(nbytes, 16) // 8 + 1, dtype=torch.float64,
                                                       device=self._buf.device)
            N.call("pyftle_pack_spline", ctypes.byref(self._desc), int(k), ptr(src), dtype_code(src.dtype),
                   arr, ptr(work), stream_handle(stream))
        else:
            N.call("pyftle_pack_snapshot", ctypes.byref(self._desc), int(k), ptr(src),
                   dtype_code(src.dtype), arr, stream_handle(stream))
        self.resident[self.slot_of(k)] = k

    def sample(self, points, time_mode=N.TIME_STEADY, snapshot=0, t=0.0):
        """Velocity at physical points (``HighPerfInterpolator.interpolate``)."""
        is_np = not isinstance(points, torch.Tensor)
        pts = to_device(points, torch.float64 if is_np else None)
        if pts.dtype not in (torch.float32, torch.float64):
            pts = pts.to(torch.float64)
        out = torch.empty_like(pts)
        N.call("pyftle_field_sample", ctypes.byref(self._desc), int(time_mode), int(snapshot),
               float(t), ptr(pts), dtype_code(pts.dtype), pts.shape[0], ptr(out), stream_handle())
        if is_np:
            return out.cpu().numpy()
        return out
