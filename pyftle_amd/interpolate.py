"""Interpolator strategies (interpolate.py:19-111, 297-543, 765-901).

``HighPerfInterpolator`` ("grid") is the device strategy: every ``update``
packs one velocity snapshot into a ring of HBM slots (HIP pack kernel) and
``interpolate`` samples it with the HIP gather kernel.  The integrators read
the same ring from inside the fused advection kernel, so in the window driver
no velocity ever returns to the host.

Differences from the reference, by design:
* ``update(velocities)`` without points reuses the stored grid geometry; the
  reference assigns a 1-D view there and its C++ then rejects it
  (interpolate.py:468-472, GridInterp.cpp:96).
* Compute precision is a constructor argument: ``"float32"`` (default, the
  fast path) or ``"float64"`` (the reference's arithmetic, operation for
  operation).
* ``cubic``, ``linear`` and ``nearest`` with a ``grid_shape`` are the SciPy
  ``RegularGridInterpolator`` strategies (``GridInterpolator``,
  interpolate.py:546-762), run by the same device kernels in their own
  interpolation modes.  Without a grid shape they are SciPy's scattered-data
  strategies, also on the device (interpolate.py:114-294):
  ``CubicInterpolator`` (Clough-Tocher, 2D), ``LinearInterpolator`` (Delaunay +
  barycentric), ``NearestNeighborInterpolator``.
"""

from __future__ import annotations

import hashlib
from abc import ABC, abstractmethod
from typing import Callable, Optional

import numpy as np
import torch

from . import _native as N
from .field import DeviceField, GridGeometry, precision_dtype


class ExecutableNotFoundError(RuntimeError):
    """Raised when an analytical interpolator has no callable ``velocity_fn``.

    The reference raises matplotlib's exception of the same name
    (interpolate.py:6,809-810); this class keeps the name without the
    matplotlib dependency.
    """


class Interpolator(ABC):
    """interpolate.py:19-111: lazy strategy, ``update`` then ``interpolate``."""

    def __init__(self):
        self.velocities = None
        self.points = None
        self.interpolator = None
        self.velocity_fn: Optional[Callable] = None
        self.grid_shape: Optional[tuple] = None

    def _initialize_interpolator(self) -> None:
        if self.velocities is None or self.points is None:
            raise ValueError("Velocities and points must be set before initialization.")
        if self.velocities.shape[-1] != self.points.shape[-1]:
            raise ValueError("Number of velocities must match the number of points.")
        raise NotImplementedError("This method should be implemented by subclasses")

    def update(self, velocities, points=None) -> None:
        self.velocities = velocities
        if points is not None:
            self.points = points
        self._initialize_interpolator()

    @abstractmethod
    def interpolate(self, new_points):
        pass


class HighPerfInterpolator(Interpolator):
    """Device grid interpolator (interpolate.py:297-543), bi/trilinear.

    Parameters
    ----------
    grid_shape : tuple
        ``(nx, ny)`` or ``(nx, ny, nz)`` as in the reference.
    precision : "float32" | "float64"
        Storage and compute precision of the device field.
    ring_slots : int
        Snapshots kept resident; the window driver raises it to the window
        length (up to ``max_ring_bytes``).
    periodic : tuple of bool, optional
        Wrap indices on these axes instead of returning zero outside the grid
        (an extension; the reference has no periodic grids).
    """

    def __init__(self, grid_shape, precision="float32", ring_slots: int = 1, periodic=None,
                 max_ring_bytes: int = 8 << 30):
        super().__init__()
        self.grid_shape = tuple(int(s) for s in grid_shape)
        self.dim = len(self.grid_shape)
        if self.dim not in (2, 3):
            raise ValueError(f"Unsupported grid dimensionality: {self.dim}")
        self.dtype = precision_dtype(precision)
        self.periodic = periodic
        self.geom: Optional[GridGeometry] = None
        self.field: Optional[DeviceField] = None
        self._ring = max(1, int(ring_slots))
        self.max_ring_bytes = int(max_ring_bytes)
        self.current = -1  # snapshot index of the last update

    # -- ring management -------------------------------------------------
    def slot_bytes(self) -> int:
        comps = 2 if self.dim == 2 else 4
        esize = 4 if self.dtype == torch.float32 else 8
        return int(np.prod(self.grid_shape)) * comps * esize

    def reserve(self, n_snapshots: int) -> int:
        """Make room for ``n_snapshots`` resident slots (bounded by memory)."""
        cap = max(1, self.max_ring_bytes // max(1, self.slot_bytes()))
        want = max(1, min(int(n_snapshots), cap))
        if want != self._ring:
            self._ring = want
            if self.geom is not None:
                self._allocate()
        return self._ring

    def begin_window(self, n_snapshots: int) -> int:
        """Start a new window: numbering restarts at snapshot 0."""
        ring = self.reserve(n_snapshots)
        self.current = -1
        return ring

    def _allocate(self) -> None:
        self.field = DeviceField(self.geom, dtype=self.dtype, layout=N.LAYOUT_SINGLE,
                                 ring_slots=self._ring, n_snapshots=1 << 30)

    def _initialize_interpolator(self) -> None:
        if self.velocities is None or self.points is None:
            raise ValueError("Velocities and points must be set before initialization.")
        geom = GridGeometry.from_reference(self.grid_shape, self.points, self.periodic)
        if self.field is None or geom != self.geom:
            self.geom = geom
            self._allocate()

    def update(self, velocities, points=None) -> None:
        """interpolate.py:446-472: pack the snapshot into the next ring slot."""
        self.velocities = velocities
        if points is not None:
            self.points = points
            self._initialize_interpolator()
        elif self.field is None:
            raise ValueError("Velocities and points must be set before initialization.")
        vel = velocities
        ngrid = int(np.prod(self.grid_shape))
        if tuple(vel.shape) != (self.dim, ngrid):
            raise ValueError(f"velocities must have shape {(self.dim, ngrid)}, got {tuple(vel.shape)}")
        self.current += 1
        self.field.pack(self.current, vel)

    def device_field(self):
        if self.field is None or self.current < 0:
            raise ValueError("Velocities and points must be set before initialization.")
        return self.field, self.current

    def interpolate(self, new_points, out=None):
        """interpolate.py:475-543: velocity ``(M, dim)`` at physical points."""
        dim = new_points.shape[1]
        if dim != self.dim:
            raise ValueError(f"Unsupported point dimension: {dim}")
        field, k = self.device_field()
        res = field.sample(new_points, time_mode=N.TIME_FROZEN, snapshot=k)
        if out is not None:
            out[...] = res
            return out
        return res


class GridInterpolator(HighPerfInterpolator):
    """SciPy ``RegularGridInterpolator`` strategies on the device (interpolate.py:546-762).

    ``method`` is ``"linear"`` (or ``"slinear"``, the same piecewise-linear
    function), ``"nearest"`` or ``"cubic"``; as in the reference each velocity
    component is reshaped to ``grid_shape`` in C order on the axes
    ``numpy.linspace(points[a].min(), points[a].max(), n_a)``, points outside the
    grid get 0 (``bounds_error=False, fill_value=0``), the upper edge belongs to
    the grid, NaN positions give NaN.  ``"cubic"`` fits the tensor-product
    not-a-knot cubic spline of every snapshot exactly (SciPy's default iterative
    solver reaches it to ~1e-5); ``"quintic"`` is not provided.

    Deliberate differences, both for updates without new points: the reference
    re-reads later 2D snapshots in Fortran order although the first one was read
    in C order, and its cubic splines stay fitted to the first snapshot
    (``RegularGridInterpolator.values`` is replaced without a refit).  Here
    every snapshot is read in C order and fitted.
    """

    VALID_METHODS = {"linear", "nearest", "slinear", "cubic", "quintic"}
    _MODES = {"linear": N.INTERP_RGI_LINEAR, "slinear": N.INTERP_RGI_LINEAR,
              "nearest": N.INTERP_RGI_NEAREST, "cubic": N.INTERP_RGI_CUBIC}

    def __init__(self, grid_shape, method: str = "linear", precision="float32", ring_slots: int = 1,
                 max_ring_bytes: int = 8 << 30):
        if method not in self.VALID_METHODS:
            raise ValueError(f"Invalid method '{method}'. Must be one of {self.VALID_METHODS}")
        if method == "quintic":
            raise NotImplementedError("GridInterpolator: quintic splines have no device implementation")
        super().__init__(grid_shape, precision=precision, ring_slots=ring_slots, max_ring_bytes=max_ring_bytes)
        self.method = method
        self.mode = self._MODES[method]
        if self.mode == N.INTERP_RGI_CUBIC and min(self.grid_shape) < 4:
            i = int(np.argmin(self.grid_shape))
            raise ValueError(f"There are {self.grid_shape[i]} points in dimension {i}, but method cubic "
                             "requires at least  4 points per dimension.")

    def slot_bytes(self) -> int:
        comps = 2 if self.dim == 2 else 4
        esize = 4 if self.dtype == torch.float32 else 8
        ext = 2 if self.mode == N.INTERP_RGI_CUBIC else 0
        return int(np.prod([n + ext for n in self.grid_shape])) * comps * esize

    def _allocate(self) -> None:
        self.field = DeviceField(self.geom, dtype=self.dtype, layout=N.LAYOUT_SINGLE,
                                 ring_slots=self._ring, n_snapshots=1 << 30, interp=self.mode)

    def _initialize_interpolator(self) -> None:
        if self.velocities is None or self.points is None:
            raise ValueError("Velocities and points must be set before initialization.")
        geom = GridGeometry.from_rgi(self.grid_shape, self.points)
        if self.field is None or geom != self.geom:
            self.geom = geom
            self._allocate()


class _ScatteredInterpolator(HighPerfInterpolator):
    """Scattered-data strategies on the device (interpolate.py:188-294): the data
    points are triangulated / bucketed once per point set (``ScatterData``, cached
    by content), each snapshot's velocities become one record per point in the
    device ring, and the kernels search and blend on the GPU."""

    _mode = N.INTERP_SCATTER_LINEAR

    def __init__(self, precision="float32", ring_slots: int = 1, max_ring_bytes: int = 8 << 30):
        Interpolator.__init__(self)
        self.dim = None
        self.dtype = precision_dtype(precision)
        self.periodic = None
        self.geom: Optional[GridGeometry] = None
        self.field: Optional[DeviceField] = None
        self._ring = max(1, int(ring_slots))
        self.max_ring_bytes = int(max_ring_bytes)
        self.current = -1
        self._scatter = None
        self._points_key = None

    def slot_bytes(self) -> int:
        if self.geom is None:
            return 1
        comps = 2 if self.dim == 2 else 4
        esize = 4 if self.dtype == torch.float32 else 8
        return self.geom.n[0] * comps * esize

    def _allocate(self) -> None:
        self.field = DeviceField(self.geom, dtype=self.dtype, layout=N.LAYOUT_SINGLE, ring_slots=self._ring,
                                 n_snapshots=1 << 30, interp=self._mode, scatter=self._scatter)

    def _initialize_interpolator(self) -> None:
        if self.velocities is None or self.points is None:
            raise ValueError("Velocities and points must be set before initialization.")
        pts = self.points.detach().cpu().numpy() if isinstance(self.points, torch.Tensor) else np.asarray(self.points)
        pts = np.ascontiguousarray(pts, dtype=np.float64)
        if pts.ndim != 2 or pts.shape[0] not in (2, 3):
            raise ValueError("points must have shape (dim, N) with dim 2 or 3")
        if np.shape(self.velocities)[-1] != pts.shape[-1]:
            raise ValueError("Number of velocities must match the number of points.")
        self._check_dim(pts.shape[0])
        key = (pts.shape, hashlib.sha1(pts.tobytes()).hexdigest())
        if key != self._points_key:
            from .scatter import ScatterData

            self._scatter = ScatterData(pts, need_simplices=self._mode != N.INTERP_SCATTER_NEAREST,
                                        cubic=self._mode == N.INTERP_SCATTER_CUBIC)
            self._points_key = key
            self.dim = pts.shape[0]
            self.geom = GridGeometry.scattered(self.dim, pts.shape[1])
            self._allocate()

    def _check_dim(self, dim: int) -> None:
        pass

    def update(self, velocities, points=None) -> None:
        """interpolate.py:76-93: a new snapshot (and point set) for the next ring slot."""
        self.velocities = velocities
        if points is not None:
            self.points = points
        if self.points is None:
            raise ValueError("Velocities and points must be set before initialization.")
        self._initialize_interpolator()
        n = self.geom.n[0]
        if tuple(np.shape(velocities)) != (self.dim, n):
            raise ValueError(f"velocities must have shape {(self.dim, n)}, got {tuple(np.shape(velocities))}")
        self.current += 1
        self._pack(self.current, velocities)

    def _pack(self, k, velocities) -> None:
        self.field.pack(k, velocities)


class CubicInterpolator(_ScatteredInterpolator):
    """``CloughTocher2DInterpolator`` (interpolate.py:114-185), 2D only: C1 piecewise
    cubic Clough-Tocher patches on the Delaunay triangulation, NaN outside the hull.
    The vertex gradients are SciPy's global curvature-minimising estimate
    (``CloughTocher2DInterpolator(tri, u + 1j v).grad``: the reference's own call,
    one host call per snapshot); locating the triangle and evaluating the patch run
    on the GPU."""

    _mode = N.INTERP_SCATTER_CUBIC

    def _check_dim(self, dim: int) -> None:
        if dim == 3:
            raise ValueError("cubic interpolator is only valid for 2D cases")

    def slot_bytes(self) -> int:
        if self.geom is None:
            return 1
        return self.geom.n[0] * 8 * (4 if self.dtype == torch.float32 else 8)

    def _pack(self, k, velocities) -> None:
        from scipy.interpolate import CloughTocher2DInterpolator

        vel = velocities.detach().cpu().numpy() if isinstance(velocities, torch.Tensor) else np.asarray(velocities)
        vel = np.asarray(vel, dtype=np.float64)
        ct = CloughTocher2DInterpolator(self._scatter.tri, vel[0] + 1j * vel[1])
        grad = np.asarray(ct.grad).reshape(vel.shape[1], 2)
        rec = np.zeros((vel.shape[1], 8))
        rec[:, 0], rec[:, 1] = vel[0], vel[1]
        rec[:, 2:4] = grad.real
        rec[:, 4:6] = grad.imag
        self.field.write_records(k, rec)


class LinearInterpolator(_ScatteredInterpolator):
    """``LinearNDInterpolator`` on the Delaunay triangulation of the data points
    (interpolate.py:188-246): barycentric blend in the containing simplex, NaN
    outside the convex hull.  The reference fills the third component in 3D only
    when exactly three points are queried (interpolate.py:244 tests the number of
    points, not the dimension) and leaves it uninitialised otherwise; here all
    components are always interpolated."""

    _mode = N.INTERP_SCATTER_LINEAR


class NearestNeighborInterpolator(_ScatteredInterpolator):
    """``NearestNDInterpolator`` (interpolate.py:249-294): the velocity of the
    nearest data point (ties: the lower point index), defined everywhere."""

    _mode = N.INTERP_SCATTER_NEAREST


class AnalyticalInterpolator(Interpolator):
    """interpolate.py:765-829: ``velocity_fn(time, points)`` with time fixed at 0.

    The device integrators call the function exactly where the reference does: at
    every integrator stage, with ``time`` (0.0) and the ``(M, dim)`` stage positions,
    here a CUDA tensor of the positions' dtype (FP64 for the reference's arrays).  It
    must return the ``(M, dim)`` velocities on that device, so it is written with
    torch operations (``torch.sin``, ``torch.exp`` ...); a function built on NumPy
    ufuncs cannot take device tensors and raises ``TypeError`` saying so.  Global
    reductions inside the function (a normalisation by the largest speed, as in the
    reference's helical-vortex example) therefore see the same point set as in the
    reference.  The HIP stage kernels (``pyftle_stage_combine``) combine the stages.

    Opt-in: ``set_grid(lower, upper, shape)`` samples the function once, at
    ``time``, onto a uniform device grid that the fused window kernel reads
    (bilinear / trilinear error O(spacing^2) of the field's second derivatives);
    a particle that leaves the sampled box makes the window raise ``ValueError``.
    """

    def __init__(self, velocity_fn: Callable):
        super().__init__()
        self.velocity_fn = velocity_fn
        self.time = 0.0
        self._grid_spec = None
        self._grid = None

    def interpolate(self, new_points):
        if not callable(self.velocity_fn):
            raise ExecutableNotFoundError("velocity_fn was not assigned properly")
        if isinstance(new_points, torch.Tensor) and new_points.is_cuda:
            return self.evaluate(new_points)
        return self.velocity_fn(self.time, new_points)

    def evaluate(self, points: torch.Tensor) -> torch.Tensor:
        """``velocity_fn(time, points)`` on a device tensor, checked and made a
        contiguous tensor of the points' dtype and shape."""
        if not callable(self.velocity_fn):
            raise ExecutableNotFoundError("velocity_fn was not assigned properly")
        try:
            v = self.velocity_fn(self.time, points)
        except TypeError as e:
            raise TypeError(
                "velocity_fn(t, points) is called with a CUDA tensor on the device path and must "
                "return one (write it with torch operations, e.g. torch.sin instead of np.sin), "
                "or sample it onto a device grid with set_grid() / AnalyticalSolver(grid=...)"
            ) from e
        if not isinstance(v, torch.Tensor):
            v = torch.as_tensor(np.asarray(v), device=points.device)
        if tuple(v.shape) != tuple(points.shape):
            raise ValueError(f"velocity_fn returned shape {tuple(v.shape)} for points of shape "
                             f"{tuple(points.shape)}")
        if v.device != points.device or v.dtype != points.dtype:
            v = v.to(device=points.device, dtype=points.dtype)
        v = v.contiguous()
        if v.data_ptr() == points.data_ptr():  # the stage buffer is rewritten before v is used
            v = v.clone()
        return v

    def update(self, velocities, points=None) -> None:
        pass

    @property
    def is_gridded(self) -> bool:
        return self._grid_spec is not None

    def set_grid(self, lower, upper, shape, precision="float32", periodic=None) -> None:
        """Opt in to evaluating the field once on a uniform device grid over
        ``[lower, upper]`` with ``shape`` nodes per axis."""
        self._grid_spec = (tuple(lower), tuple(upper), tuple(int(s) for s in shape), precision, periodic)
        self._grid = None

    def gridded(self) -> HighPerfInterpolator:
        if self._grid is not None:
            return self._grid
        if self._grid_spec is None:
            raise ValueError("AnalyticalInterpolator.gridded() needs set_grid() first")
        lower, upper, shape, precision, periodic = self._grid_spec
        axes = [np.linspace(lower[a], upper[a], shape[a]) for a in range(len(shape))]
        if len(shape) == 2:
            X, Y = np.meshgrid(*axes)  # x-fastest flat order (interpolate.py:363)
            pts = np.stack([X.ravel(), Y.ravel()], axis=1)
        else:
            X, Y, Z = np.meshgrid(*axes)  # 'xy' meshgrid, flat C order (interpolate.py:383-385)
            pts = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
        dev_pts = torch.as_tensor(pts, device="cuda")
        try:
            vel = self.evaluate(dev_pts).to(torch.float64)
        except TypeError:
            # a NumPy-only function: sampling the grid is ingest (like reading a snapshot
            # file), done once on the host; the window itself runs on the device
            vel = torch.as_tensor(np.asarray(self.velocity_fn(self.time, pts), dtype=np.float64),
                                  device="cuda")
        g = HighPerfInterpolator(shape, precision=precision, periodic=periodic)
        g.update(vel.t().contiguous(), dev_pts.t().contiguous())
        self._grid = g
        return g

    def check_inside_grid(self, positions: torch.Tensor) -> None:
        """Raises ``ValueError`` if a particle lies outside the sampled box (where the
        gridded field is zero and the particle would have stopped)."""
        lower, upper, shape, _, periodic = self._grid_spec
        lo = torch.as_tensor(lower, dtype=positions.dtype, device=positions.device)
        hi = torch.as_tensor(upper, dtype=positions.dtype, device=positions.device)
        out = (positions < lo) | (positions > hi)
        if periodic is not None:
            out &= ~torch.as_tensor([bool(p) for p in periodic], device=positions.device)
        n_out = int(out.any(dim=1).sum().item())
        if n_out:
            raise ValueError(f"{n_out} particles left the sampled box [{lower}, {upper}] of the gridded "
                             "analytical field; enlarge the grid passed to set_grid()")


def create_interpolator(interpolation_type: str, grid_shape=None, velocity_fn=None, **kwargs) -> Interpolator:
    """interpolate.py:832-901 factory (case-insensitive names)."""
    interpolation_type = interpolation_type.lower()
    names = ["cubic", "linear", "nearest", "analytical", "grid"]
    if interpolation_type not in names:
        raise ValueError(f"Invalid interpolation type '{interpolation_type}'. Choose from {names}.")
    if interpolation_type == "analytical":
        if not callable(velocity_fn):
            raise ExecutableNotFoundError("velocity_fn was not assigned properly")
        return AnalyticalInterpolator(velocity_fn)
    if interpolation_type == "grid":
        if grid_shape is None:
            raise ValueError("the 'grid' interpolator needs grid_shape")
        return HighPerfInterpolator(grid_shape, **kwargs)
    if grid_shape is not None:  # interpolate.py:895-896: RegularGridInterpolator strategies
        return GridInterpolator(grid_shape, interpolation_type, **kwargs)
    if interpolation_type == "linear":
        return LinearInterpolator(**kwargs)
    if interpolation_type == "nearest":
        return NearestNeighborInterpolator(**kwargs)
    return CubicInterpolator(**kwargs)
